// Host side of the MI355X backend: see phase.hpp.
#include "phase.hpp"

#include <algorithm>
#include <cctype>
#include <cinttypes>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>

namespace phase {

// ===========================================================================
// Errors
// ===========================================================================
[[noreturn]] void raise_status(int code, const std::string &context) {
  std::string msg = context;
  const char *le = ph_last_error();
  if (le && *le) msg += ": " + std::string(le);
  switch (code) {
    case PH_E_INDEX: throw IndexError(msg);
    case PH_E_SHAPE: throw ShapeError(msg);
    case PH_E_DIMENSION: throw DimensionError(msg);
    case PH_E_ARGUMENT: throw ArgumentError(msg);
    case PH_E_OVERFLOW: throw OverflowError(msg);
    case PH_E_DIVZERO: throw DivisionByZeroError(msg);
    default: throw DeviceError(msg);
  }
}

void check(int code, const char *context) {
  if (code != PH_OK) raise_status(code, context);
}

static void check_flags(int32_t flags, const char *what) {
  if (flags & PH_FLAG_NEGEXP)
    throw ArgumentError(std::string(what) + ": Cannot raise an integer to a negative integer power, use floats for that");
  if (flags & PH_FLAG_DIVZERO) throw DivisionByZeroError(std::string(what) + ": Division by 0");
  if (flags & PH_FLAG_OVERFLOW) throw OverflowError(std::string(what) + ": Arithmetic overflow");
}

static bool g_inited = false;

void init(int device) {
  if (g_inited) return;
  check(ph_init(device), "ph_init");
  g_inited = true;
}

bool device_available() {
  int n = 0;
  if (ph_device_count(&n) != PH_OK) return false;
  return n > 0;
}

// ===========================================================================
// DType / Scalar
// ===========================================================================
const char *dtype_name(DType d) {
  switch (d) {
    case DType::I32: return "Int32";
    case DType::F32: return "Float32";
    case DType::F64: return "Float64";
    case DType::BF16: return "BFloat16";
    case DType::Bool: return "Bool";
  }
  return "?";
}

size_t dtype_size(DType d) {
  switch (d) {
    case DType::I32: return 4;
    case DType::F32: return 4;
    case DType::F64: return 8;
    case DType::BF16: return 2;
    case DType::Bool: return 1;
  }
  return 0;
}

DType compute_dtype(DType s) { return s == DType::BF16 ? DType::F32 : s; }

Scalar Scalar::i32(int64_t v) {
  if (v < INT32_MIN || v > INT32_MAX) throw OverflowError("Int32 literal out of range: " + std::to_string(v));
  Scalar s;
  s.dt = DType::I32;
  s.i = v;
  return s;
}
Scalar Scalar::f32(double v) {
  Scalar s;
  s.dt = DType::F32;
  s.f = (double)(float)v;
  return s;
}
Scalar Scalar::f64(double v) {
  Scalar s;
  s.dt = DType::F64;
  s.f = v;
  return s;
}
Scalar Scalar::boolean(bool v) {
  Scalar s;
  s.dt = DType::Bool;
  s.i = v ? 1 : 0;
  return s;
}
std::string Scalar::to_s() const {
  switch (dt) {
    case DType::I32: return std::to_string(i);
    case DType::Bool: return i ? "true" : "false";
    default: {
      char b[64];
      std::snprintf(b, sizeof(b), "%.17g", f);
      return b;
    }
  }
}

// bits of a scalar stored as `dt`
static void scalar_bits(const Scalar &s, DType dt, void *out) {
  switch (dt) {
    case DType::I32: {
      int32_t v = (int32_t)(s.dt == DType::I32 || s.dt == DType::Bool ? s.i : (int64_t)s.f);
      std::memcpy(out, &v, 4);
      break;
    }
    case DType::F32: {
      float v = (float)(s.dt == DType::I32 || s.dt == DType::Bool ? (double)s.i : s.f);
      std::memcpy(out, &v, 4);
      break;
    }
    case DType::F64: {
      double v = s.dt == DType::I32 || s.dt == DType::Bool ? (double)s.i : s.f;
      std::memcpy(out, &v, 8);
      break;
    }
    case DType::BF16: {
      float f = (float)(s.dt == DType::I32 || s.dt == DType::Bool ? (double)s.i : s.f);
      uint32_t u;
      std::memcpy(&u, &f, 4);
      u += 0x7FFFu + ((u >> 16) & 1u);
      uint16_t h = (uint16_t)(u >> 16);
      std::memcpy(out, &h, 2);
      break;
    }
    case DType::Bool: {
      uint8_t v = (s.dt == DType::Bool || s.dt == DType::I32) ? (s.i != 0) : (s.f != 0.0);
      std::memcpy(out, &v, 1);
      break;
    }
  }
}

static Scalar scalar_from_bits(DType dt, const void *p) {
  switch (dt) {
    case DType::I32: { int32_t v; std::memcpy(&v, p, 4); return Scalar::i32(v); }
    case DType::F32: { float v; std::memcpy(&v, p, 4); return Scalar::f32(v); }
    case DType::F64: { double v; std::memcpy(&v, p, 8); return Scalar::f64(v); }
    case DType::BF16: {
      uint16_t h; std::memcpy(&h, p, 2);
      uint32_t u = (uint32_t)h << 16; float f; std::memcpy(&f, &u, 4);
      return Scalar::f32(f);
    }
    case DType::Bool: { uint8_t v; std::memcpy(&v, p, 1); return Scalar::boolean(v != 0); }
  }
  return Scalar();
}

std::string shape_to_s(const Shape &s) {
  std::string r = "[";
  for (size_t i = 0; i < s.size(); ++i) r += (i ? ", " : "") + std::to_string(s[i]);
  return r + "]";
}

// ===========================================================================
// Range literal parsing (Crystal source forms)
// ===========================================================================
namespace {

struct RNode {  // a parsed Crystal range expression
  bool is_range = false;
  bool has_value = false;  // Int (when !is_range)
  int64_t value = 0;
  std::shared_ptr<RNode> begin, end;
  bool exclusive = false;
};

struct RParser {
  std::string s;
  size_t p = 0;
  void ws() {
    while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
  }
  [[noreturn]] void fail(const std::string &why) {
    throw ArgumentError("poorly formatted range literal '" + s + "': " + why +
                        " (should be a..c, a..b..c, (a..b)..c, or a..(b..c) where a,c are Int | Nil, b is Int)");
  }
  bool at_dots(bool *excl) {
    ws();
    if (s.compare(p, 3, "...") == 0) {
      *excl = true;
      return true;
    }
    if (s.compare(p, 2, "..") == 0) {
      *excl = false;
      return true;
    }
    return false;
  }
  std::shared_ptr<RNode> atom() {  // Int, (range), or empty
    ws();
    auto n = std::make_shared<RNode>();
    if (p < s.size() && s[p] == '(') {
      ++p;
      n = expr();
      ws();
      if (p >= s.size() || s[p] != ')') fail("missing ')'");
      ++p;
      return n;
    }
    size_t b = p;
    if (p < s.size() && (s[p] == '-' || s[p] == '+')) ++p;
    while (p < s.size() && (std::isdigit((unsigned char)s[p]) || s[p] == '_')) ++p;
    if (p == b) return n;  // empty operand (nil)
    std::string num;
    for (size_t i = b; i < p; ++i)
      if (s[i] != '_') num += s[i];
    if (num == "-" || num == "+") fail("sign without digits");
    errno = 0;
    long long v = std::strtoll(num.c_str(), nullptr, 10);
    if (errno) fail("integer out of range");
    n->has_value = true;
    n->value = v;
    return n;
  }
  std::shared_ptr<RNode> expr() {
    auto left = atom();
    bool excl;
    while (at_dots(&excl)) {
      p += excl ? 3 : 2;
      auto right = atom();
      auto r = std::make_shared<RNode>();
      r->is_range = true;
      r->begin = left;
      r->end = right;
      r->exclusive = excl;
      left = r;
    }
    return left;
  }
};

}  // namespace

RangeLiteral RangeLiteral::of_index(int64_t i) {
  RangeLiteral r;
  r.is_index = true;
  r.index = i;
  r.text = std::to_string(i);
  return r;
}

RangeLiteral RangeLiteral::full() { return parse(".."); }

RangeLiteral RangeLiteral::parse(const std::string &text) {
  RangeLiteral r;
  r.text = text;
  // Steppable::StepIterator form: a.step(by: b, to: c[, exclusive: true])
  size_t sp = text.find(".step(");
  if (sp != std::string::npos) {
    auto num_after = [&](const std::string &key, bool *found) -> int64_t {
      size_t k = text.find(key);
      *found = k != std::string::npos;
      if (!*found) return 0;
      return std::strtoll(text.c_str() + k + key.size(), nullptr, 10);
    };
    bool fb, ft;
    std::string head = text.substr(0, sp);
    while (!head.empty() && (head[0] == '(' || std::isspace((unsigned char)head[0]))) head.erase(0, 1);
    while (!head.empty() && (head.back() == ')' || std::isspace((unsigned char)head.back()))) head.pop_back();
    r.first = std::strtoll(head.c_str(), nullptr, 10);
    r.step = num_after("by:", &fb);
    int64_t to = num_after("to:", &ft);
    if (!fb) throw ArgumentError("step literal without by: '" + text + "'");
    if (ft) r.last = to;
    r.exclusive = text.find("exclusive: true") != std::string::npos ||
                  text.find("exclusive:true") != std::string::npos;
    if (*r.step == 0) throw ArgumentError("Zero step size");
    return r;
  }
  RParser ps;
  ps.s = text;
  auto n = ps.expr();
  ps.ws();
  if (ps.p != text.size()) ps.fail("trailing characters");
  if (!n->is_range) {
    if (!n->has_value) ps.fail("empty literal");
    r.is_index = true;
    r.index = n->value;
    return r;
  }
  // parse_range (range_syntax.cr:41-59)
  auto intish = [](const std::shared_ptr<RNode> &x) { return !x->is_range; };
  auto val = [](const std::shared_ptr<RNode> &x) -> std::optional<int64_t> {
    if (x->has_value) return x->value;
    return std::nullopt;
  };
  if (n->begin->is_range) {
    if (!intish(n->end) || !intish(n->begin->begin) || !intish(n->begin->end)) ps.fail("nested ranges");
    r.first = val(n->begin->begin);
    r.step = val(n->begin->end);
    r.last = val(n->end);
    r.exclusive = n->exclusive;
  } else if (n->end->is_range) {
    if (!intish(n->end->begin) || !intish(n->end->end)) ps.fail("nested ranges");
    r.first = val(n->begin);
    r.step = val(n->end->begin);
    r.last = val(n->end->end);
    r.exclusive = n->end->exclusive;
  } else {
    r.first = val(n->begin);
    r.last = val(n->end);
    r.exclusive = n->exclusive;
  }
  return r;
}

// ===========================================================================
// RangeSyntax / CoordUtil / ShapeUtil
// ===========================================================================
static int64_t sign_of(int64_t v) { return (v > 0) - (v < 0); }
static int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}
static int64_t floormod(int64_t a, int64_t b) {
  int64_t m = a % b;
  if (m != 0 && ((m < 0) != (b < 0))) m += b;
  return m;
}

namespace RangeSyntax {

int64_t get_size(int64_t first, int64_t last, int64_t step) {
  int64_t cmp = (last > first) - (last < first);
  if (last != first && sign_of(step) != cmp)
    throw IndexError("Could not get size of range - step direction disagrees with first and last.");
  if (step == 0) throw DivisionByZeroError("Division by 0");
  if (last >= first) return floordiv(last - first, step) + 1;
  return floordiv(first - last, -step) + 1;
}

static int64_t canon_unsafe(int64_t i, int64_t size) { return i < 0 ? size + i : i; }

CanonicalRange infer_range(const RangeLiteral &lit, int64_t bound) {
  if (lit.is_index) {
    int64_t c = canon_unsafe(lit.index, bound);
    return {c, 1, c, 1};
  }
  int64_t step, first, temp_last;
  if (!lit.step) {
    first = lit.first ? canon_unsafe(*lit.first, bound) : 0;
    temp_last = lit.last ? canon_unsafe(*lit.last, bound) : bound - 1;
    step = temp_last >= first ? 1 : -1;
  } else {
    step = *lit.step;
    first = lit.first ? canon_unsafe(*lit.first, bound) : (step > 0 ? 0 : bound - 1);
    temp_last = lit.last ? canon_unsafe(*lit.last, bound) : (step > 0 ? bound - 1 : 0);
  }
  if (lit.last && lit.exclusive) {
    if (temp_last == first) return {0, 0, 0, 0};
    temp_last -= sign_of(step);
  }
  if (first < 0 || temp_last < 0)
    throw IndexError("Invalid index: At least one endpoint of " + lit.to_s() +
                     " is negative after canonicalization.");
  int64_t size;
  try {
    size = get_size(first, temp_last, step);
  } catch (const IndexError &) {
    throw IndexError("Could not canonicalize range: Conflict between implicit direction of " + lit.to_s() +
                     " and provided step " + std::to_string(step));
  }
  return {first, step, first + step * (size - 1), size};
}

CanonicalRange canonicalize_range(const RangeLiteral &lit, int64_t bound) {
  CanonicalRange r = infer_range(lit, bound);
  bool valid = r.first >= 0 && r.first < bound && r.last >= 0 && r.last < bound;
  if (!valid)
    throw IndexError("Could not canonicalize range: " + lit.to_s() +
                     " is not a sensible index range for axis of length " + std::to_string(bound) + ".");
  return r;
}

bool bounded(const RangeLiteral &lit) {
  if (lit.is_index) return true;
  if (!lit.step || *lit.step >= 0) return lit.last.has_value();
  return lit.first.has_value();
}

void ensure_nonnegative(const RangeLiteral &lit) {
  bool neg = lit.is_index ? lit.index < 0 : ((lit.first && *lit.first < 0) || (lit.last && *lit.last < 0));
  if (neg) throw IndexError("Negative indices have no meaning when a bounding shape is not provided.");
}

}  // namespace RangeSyntax

namespace CoordUtil {
bool has_index(int64_t index, int64_t size) { return index < size && index >= -size; }
int64_t canonicalize_index(int64_t index, int64_t size) {
  if (!has_index(index, size))
    throw IndexError("Could not canonicalize index: " + std::to_string(index) +
                     " is not a valid index for an axis of length " + std::to_string(size) + ".");
  return index < 0 ? size + index : index;
}
Coord canonicalize_coord(const Coord &coord, const Shape &shape) {
  if (coord.size() != shape.size())
    throw DimensionError("Could not canonicalize coordinate " + shape_to_s(coord) + " to fit in shape " +
                         shape_to_s(shape) + ": shape has " + std::to_string(shape.size()) +
                         " dimensions, while coord has " + std::to_string(coord.size()) + ".");
  Coord r(coord.size());
  for (size_t i = 0; i < coord.size(); ++i) r[i] = canonicalize_index(coord[i], shape[i]);
  return r;
}
}  // namespace CoordUtil

namespace ShapeUtil {
bool compatible_shapes(const Shape &a, const Shape &b) {
  if (a.empty() || b.empty()) return a == b;
  const Shape &larger = a.size() > b.size() ? a : b;
  size_t shared = std::min(a.size(), b.size());
  for (size_t i = 0; i < shared; ++i)
    if (a[i] != b[i]) return false;
  for (size_t i = shared; i < larger.size(); ++i)
    if (larger[i] != 1) return false;
  return true;
}
int64_t shape_to_size(const Shape &shape) {
  if (shape.empty()) return 0;
  int64_t n = 1;
  for (int64_t s : shape) n *= s;
  return n;
}
}  // namespace ShapeUtil

// ===========================================================================
// IndexRegion
// ===========================================================================
void IndexRegion::compute_reduced() {
  if (drop) {
    reduced_shape.clear();
    for (size_t i = 0; i < proper_shape.size(); ++i)
      if (!degeneracy[i]) reduced_shape.push_back(proper_shape[i]);
    if (reduced_shape.empty()) reduced_shape.push_back(ShapeUtil::shape_to_size(proper_shape));
  } else {
    reduced_shape = proper_shape;
  }
}

IndexRegion IndexRegion::from_literal(const RegionLiteral &lit, const Shape &bound, bool drop) {
  if (lit.size() > bound.size())
    throw DimensionError("The region literal had more dimensions than its bound shape " + shape_to_s(bound));
  IndexRegion r;
  const size_t n = bound.size();
  r.first.assign(n, 0);
  r.step.assign(n, 1);
  r.last.assign(n, 0);
  r.proper_shape.assign(n, 0);
  r.degeneracy.assign(n, false);
  r.drop = drop;
  for (size_t i = 0; i < lit.size(); ++i) {
    CanonicalRange c = RangeSyntax::canonicalize_range(lit[i], bound[i]);
    r.first[i] = c.first;
    r.step[i] = c.step;
    r.last[i] = c.last;
    r.proper_shape[i] = c.size;
    if (lit[i].is_index) r.degeneracy[i] = drop;
  }
  for (size_t a = lit.size(); a < n; ++a) {
    r.last[a] = bound[a] - 1;
    r.proper_shape[a] = bound[a];
  }
  r.compute_reduced();
  return r;
}

IndexRegion IndexRegion::absolute(const RegionLiteral &lit, bool drop) {
  IndexRegion r;
  const size_t n = lit.size();
  r.first.assign(n, 0);
  r.step.assign(n, 0);
  r.last.assign(n, 0);
  r.proper_shape.assign(n, 0);
  r.degeneracy.assign(n, false);
  r.drop = drop;
  for (size_t i = 0; i < n; ++i) {
    RangeSyntax::ensure_nonnegative(lit[i]);
    if (!RangeSyntax::bounded(lit[i]))
      throw PhaseError("Cannot create IndexRegion without an explicit upper bound unless you provide a bounding shape");
    if (lit[i].is_index) r.degeneracy[i] = true;
    CanonicalRange c = RangeSyntax::infer_range(lit[i], 0);
    r.first[i] = c.first;
    r.step[i] = c.step;
    r.last[i] = c.last;
    r.proper_shape[i] = c.size;
  }
  r.compute_reduced();
  return r;
}

IndexRegion IndexRegion::trimmed(const RegionLiteral &lit, const Shape *bound, bool drop, const Shape &trim_to) {
  IndexRegion r;
  const size_t n = trim_to.size();
  r.first.assign(n, 0);
  r.step.assign(n, 1);
  r.last.assign(n, 0);
  r.proper_shape.assign(n, 0);
  r.degeneracy.assign(n, false);
  r.drop = drop;
  const Shape &bs = bound ? *bound : trim_to;
  for (size_t i = 0; i < lit.size(); ++i) {
    if (!bound) RangeSyntax::ensure_nonnegative(lit[i]);
    CanonicalRange c = RangeSyntax::infer_range(lit[i], bs.at(i));
    r.first[i] = c.first;
    r.step[i] = c.step;
    r.last[i] = c.last;
    r.proper_shape[i] = c.size;
    if (lit[i].is_index) r.degeneracy[i] = drop;
  }
  for (size_t a = lit.size(); a < bs.size() && a < n; ++a) {
    r.last[a] = bs[a] - 1;
    r.proper_shape[a] = bs[a];
  }
  r.compute_reduced();
  return r.trim(trim_to);
}

IndexRegion IndexRegion::cover(const Shape &shape, bool drop, const std::vector<bool> *degeneracy) {
  IndexRegion r;
  const size_t n = shape.size();
  r.first.assign(n, 0);
  r.step.resize(n);
  r.last.resize(n);
  for (size_t i = 0; i < n; ++i) {
    r.step[i] = shape[i] == 0 ? 0 : 1;
    r.last[i] = std::max<int64_t>(0, shape[i] - 1);
  }
  r.proper_shape = shape;
  r.degeneracy = degeneracy ? *degeneracy : std::vector<bool>(n, false);
  r.drop = drop;
  r.compute_reduced();
  return r;
}

IndexRegion IndexRegion::fitted_copy(const IndexRegion &src, const Shape &bound) {
  if (src.fits_in(bound)) return src;
  throw ShapeError("Region " + src.to_s() + " does not fit inside " + shape_to_s(bound));
}

IndexRegion &IndexRegion::trim(const Shape &bound) {
  if ((int)bound.size() != proper_dimensions()) throw DimensionError("invalid error :)");
  for (size_t a = 0; a < bound.size(); ++a) {
    const int64_t nb = bound[a];
    int64_t &f = first[a], &s = step[a], &l = last[a], &sz = proper_shape[a];
    if (f >= nb) {
      if (l >= nb) {
        f = 0; s = 0; l = 0; sz = 0;
      } else if (s < 0) {
        const int64_t as = -s;
        int64_t span = (nb - 1) - l;
        const int64_t size = floordiv(span, as) + 1;
        span -= floormod(span, as);
        f = l + span;
        sz = size;
      }
    } else if (s > 0 && l >= nb) {
      int64_t span = (nb - 1) - f;
      span -= floormod(span, s);
      sz = floordiv(span, s) + 1;
      l = f + span;
    }
  }
  compute_reduced();
  return *this;
}

IndexRegion &IndexRegion::reverse() {
  std::swap(first, last);
  for (auto &s : step) s = -s;
  return *this;
}

IndexRegion &IndexRegion::unsafe_translate(const std::vector<int64_t> &offset) {
  for (size_t a = 0; a < offset.size(); ++a) {
    first[a] += offset[a];
    last[a] += offset[a];
  }
  return *this;
}

IndexRegion &IndexRegion::translate(const std::vector<int64_t> &offset) {
  for (size_t a = 0; a < offset.size(); ++a)
    if (offset[a] < 0 && (first[a] < -offset[a] || last[a] < -offset[a]))
      throw IndexError("Can't translate to negative indices");
  return unsafe_translate(offset);
}

bool IndexRegion::includes(const Coord &coord) const {
  if ((int)coord.size() != proper_dimensions()) return false;
  for (size_t i = 0; i < coord.size(); ++i) {
    int64_t lo = step[i] > 0 ? first[i] : last[i];
    int64_t hi = step[i] > 0 ? last[i] : first[i];
    if (coord[i] < lo || coord[i] > hi) return false;
    if (step[i] == 0) throw DivisionByZeroError("Division by 0");
    if (floormod(coord[i] - first[i], step[i]) != 0) return false;
  }
  return true;
}

bool IndexRegion::fits_in(const Shape &bound) const {
  if ((int)bound.size() != proper_dimensions())
    throw DimensionError("The bound shape provided had a different number of dimensions than this IndexRegion, so fits_in? is meaningless.");
  for (size_t i = 0; i < bound.size(); ++i)
    if (bound[i] <= std::max(first[i], last[i])) return false;
  return true;
}

Coord IndexRegion::local_to_absolute_unsafe(const Coord &local) const {
  Coord out(proper_shape.size());
  if (drop) {
    size_t la = 0;
    for (size_t i = 0; i < out.size(); ++i) {
      if (degeneracy[i]) out[i] = first[i];
      else out[i] = first[i] + local.at(la++) * step[i];
    }
  } else {
    for (size_t i = 0; i < out.size(); ++i) out[i] = first[i] + local.at(i) * step[i];
  }
  return out;
}

Coord IndexRegion::local_to_absolute(const Coord &local) const {
  return local_to_absolute_unsafe(CoordUtil::canonicalize_coord(local, reduced_shape));
}

Coord IndexRegion::absolute_to_local_unsafe(const Coord &abs) const {
  Coord local(abs.size());
  for (size_t i = 0; i < abs.size(); ++i) local[i] = floordiv(abs[i] - first[i], step[i]);
  if (!drop) return local;
  Coord r;
  for (size_t i = 0; i < local.size(); ++i)
    if (!degeneracy[i]) r.push_back(local[i]);
  if (r.empty()) r.push_back(0);
  return r;
}

Coord IndexRegion::absolute_to_local(const Coord &abs) const {
  if (!includes(abs)) throw IndexError("Could not convert coordinate: " + shape_to_s(abs) + " does not exist in region " + to_s());
  return absolute_to_local_unsafe(abs);
}

std::string IndexRegion::to_s() const {
  std::string r = "IndexRegion[";
  for (size_t i = 0; i < first.size(); ++i) {
    if (i) r += ", ";
    if (degeneracy[i]) r += std::to_string(first[i]);
    else if (step[i] == 1 || step[i] == -1) r += std::to_string(first[i]) + ".." + std::to_string(last[i]);
    else r += std::to_string(first[i]) + ".." + std::to_string(step[i]) + ".." + std::to_string(last[i]);
  }
  return r + "]";
}

bool IndexRegion::operator==(const IndexRegion &o) const {
  return first == o.first && step == o.step && last == o.last && degeneracy == o.degeneracy && drop == o.drop;
}

// ===========================================================================
// Expressions
// ===========================================================================
struct ExprNode {
  enum Kind { Input, Const, Op, Index } kind = Const;
  int op = 0;          // ph_opcode
  DType dt = DType::I32;      // result dtype
  DType opdt = DType::I32;    // dtype the operation computes in
  int input = 0;
  DType storage = DType::I32;
  Scalar value;
  std::shared_ptr<const ExprNode> a, b, c;
  int nargs = 0;
};

static std::shared_ptr<const ExprNode> mk(ExprNode n) { return std::make_shared<const ExprNode>(std::move(n)); }

Expr Expr::input(int k, DType storage) {
  ExprNode n;
  n.kind = ExprNode::Input;
  n.input = k;
  n.storage = storage;
  n.dt = compute_dtype(storage);
  return Expr{mk(n)};
}

Expr Expr::constant(const Scalar &s) {
  ExprNode n;
  n.kind = ExprNode::Const;
  n.value = s;
  n.dt = s.dt;
  return Expr{mk(n)};
}

Expr Expr::index() {
  ExprNode n;
  n.kind = ExprNode::Index;
  n.dt = DType::I32;
  return Expr{mk(n)};
}

DType Expr::dtype() const { return node->dt; }

static Expr cast_to(const Expr &e, DType dt) {
  if (e.dtype() == dt) return e;
  if (e.node->kind == ExprNode::Const) {  // fold literal conversions
    Scalar s = e.node->value;
    Scalar r;
    if (dt == DType::F64) r = Scalar::f64(s.dt == DType::I32 || s.dt == DType::Bool ? (double)s.i : s.f);
    else if (dt == DType::F32) r = Scalar::f32(s.dt == DType::I32 || s.dt == DType::Bool ? (double)s.i : s.f);
    else if (dt == DType::I32) r = Scalar::i32(s.dt == DType::I32 || s.dt == DType::Bool ? s.i : (int64_t)s.f);
    else r = Scalar::boolean(s.dt == DType::I32 || s.dt == DType::Bool ? s.i != 0 : s.f != 0.0);
    return Expr::constant(r);
  }
  ExprNode n;
  n.kind = ExprNode::Op;
  n.op = PH_OP_CAST;
  n.dt = dt;
  n.opdt = e.dtype();
  n.a = e.node;
  n.nargs = 1;
  return Expr{mk(n)};
}

static bool is_float(DType d) { return d == DType::F32 || d == DType::F64; }

// S1: Int32 op Int32 -> Int32; any Float64 -> Float64; otherwise Float32.
static DType arith_type(DType a, DType b) {
  if (a == DType::I32 && b == DType::I32) return DType::I32;
  if (a == DType::F64 || b == DType::F64) return DType::F64;
  return DType::F32;
}

static Expr op_node(int op, DType dt, DType opdt, const Expr &a, const Expr *b, const Expr *c = nullptr) {
  ExprNode n;
  n.kind = ExprNode::Op;
  n.op = op;
  n.dt = dt;
  n.opdt = opdt;
  n.a = a.node;
  n.nargs = 1;
  if (b) {
    n.b = b->node;
    n.nargs = 2;
  }
  if (c) {
    n.c = c->node;
    n.nargs = 3;
  }
  return Expr{mk(n)};
}

static void no_bool(const std::string &op, const Expr &a, const Expr &b) {
  if (a.dtype() == DType::Bool || b.dtype() == DType::Bool)
    throw ArgumentError("undefined method '" + op + "' for Bool");
}

Expr Expr::binary(const std::string &op, const Expr &a, const Expr &b) {
  if (op == "+" || op == "-" || op == "*") {
    no_bool(op, a, b);
    DType t = arith_type(a.dtype(), b.dtype());
    int code = op == "+" ? PH_OP_ADD : op == "-" ? PH_OP_SUB : PH_OP_MUL;
    Expr x = cast_to(a, t), y = cast_to(b, t);
    return op_node(code, t, t, x, &y);
  }
  if (op == "&+" || op == "&-" || op == "&*" || op == "&**") {
    if (a.dtype() != DType::I32 || b.dtype() != DType::I32)
      throw ArgumentError("wrapping operator '" + op + "' needs Int32 operands");
    int code = op == "&+" ? PH_OP_WADD : op == "&-" ? PH_OP_WSUB : op == "&*" ? PH_OP_WMUL : PH_OP_WIPOW;
    return op_node(code, DType::I32, DType::I32, a, &b);
  }
  if (op == "/") {
    no_bool(op, a, b);
    DType t = arith_type(a.dtype(), b.dtype());
    if (t == DType::I32) t = DType::F64;  // Int32 / Int32 -> Float64
    Expr x = cast_to(a, t), y = cast_to(b, t);
    return op_node(PH_OP_DIV, t, t, x, &y);
  }
  if (op == "//" || op == "%") {
    no_bool(op, a, b);
    DType t = arith_type(a.dtype(), b.dtype());
    Expr x = cast_to(a, t), y = cast_to(b, t);
    return op_node(op == "//" ? PH_OP_FLOORDIV : PH_OP_MOD, t, t, x, &y);
  }
  if (op == "**") {
    no_bool(op, a, b);
    if (a.dtype() == DType::I32 && b.dtype() == DType::I32) return op_node(PH_OP_IPOW, DType::I32, DType::I32, a, &b);
    if (is_float(a.dtype()) && b.dtype() == DType::I32) return op_node(PH_OP_POWI, a.dtype(), a.dtype(), a, &b);
    DType t = (a.dtype() == DType::F32 && b.dtype() == DType::F32) ? DType::F32 : DType::F64;
    Expr x = cast_to(a, t), y = cast_to(b, t);
    return op_node(PH_OP_POW, t, t, x, &y);
  }
  if (op == "&" || op == "|" || op == "^") {
    int code = op == "&" ? PH_OP_AND : op == "|" ? PH_OP_OR : PH_OP_XOR;
    if (a.dtype() == DType::I32 && b.dtype() == DType::I32) return op_node(code, DType::I32, DType::I32, a, &b);
    if (a.dtype() == DType::Bool && b.dtype() == DType::Bool) return op_node(code, DType::Bool, DType::Bool, a, &b);
    throw ArgumentError("undefined method '" + op + "' for these operand types");
  }
  if (op == ">" || op == "<" || op == ">=" || op == "<=" || op == "==" || op == "!=" || op == "<=>") {
    DType t;
    if (a.dtype() == DType::Bool || b.dtype() == DType::Bool) {
      if (!(a.dtype() == DType::Bool && b.dtype() == DType::Bool) || !(op == "==" || op == "!="))
        throw ArgumentError("undefined method '" + op + "' for Bool");
      t = DType::Bool;
    } else {
      t = (is_float(a.dtype()) || is_float(b.dtype())) ? DType::F64 : DType::I32;
    }
    Expr x = cast_to(a, t), y = cast_to(b, t);
    static const std::map<std::string, int> codes = {{">", PH_OP_GT}, {"<", PH_OP_LT}, {">=", PH_OP_GE},
                                                     {"<=", PH_OP_LE}, {"==", PH_OP_EQ}, {"!=", PH_OP_NE},
                                                     {"<=>", PH_OP_CMP}};
    int code = codes.at(op);
    return op_node(code, code == PH_OP_CMP ? DType::I32 : DType::Bool, t, x, &y);
  }
  throw ArgumentError("unknown operator '" + op + "'");
}

Expr Expr::unary(const std::string &op, const Expr &a) {
  if (op == "+") return a;
  if (op == "-") {
    if (a.dtype() == DType::Bool) throw ArgumentError("undefined method '-' for Bool");
    return op_node(PH_OP_NEG, a.dtype(), a.dtype(), a, nullptr);
  }
  if (op == "~") {
    if (a.dtype() != DType::I32) throw ArgumentError("undefined method '~'");
    return op_node(PH_OP_BNOT, DType::I32, DType::I32, a, nullptr);
  }
  if (op == "abs") {
    if (a.dtype() == DType::Bool) throw ArgumentError("undefined method 'abs' for Bool");
    return op_node(PH_OP_ABS, a.dtype(), a.dtype(), a, nullptr);
  }
  if (op == "sqrt") {  // Math.sqrt: Int -> Float64
    Expr x = a.dtype() == DType::I32 ? cast_to(a, DType::F64) : a;
    if (!is_float(x.dtype())) throw ArgumentError("sqrt of Bool");
    return op_node(PH_OP_SQRT, x.dtype(), x.dtype(), x, nullptr);
  }
  if (op == "to_f32") return cast_to(a, DType::F32);
  if (op == "to_f64" || op == "to_f") return cast_to(a, DType::F64);
  if (op == "to_i32" || op == "to_i") {
    if (a.dtype() != DType::I32 && a.dtype() != DType::Bool)
      throw ArgumentError("to_i32 of a float is not supported on the device path");
    return cast_to(a, DType::I32);
  }
  throw ArgumentError("unknown unary operator '" + op + "'");
}

Expr Expr::select(const Expr &cond, const Expr &a, const Expr &b) {
  if (cond.dtype() != DType::Bool) throw ArgumentError("select condition must be Bool");
  DType t = a.dtype() == b.dtype() ? a.dtype() : arith_type(a.dtype(), b.dtype());
  Expr x = cast_to(a, t), y = cast_to(b, t);
  return op_node(PH_OP_SELECT, t, t, x, &y, &cond);
}

static std::shared_ptr<const ExprNode> subst(const std::shared_ptr<const ExprNode> &n,
                                             const std::vector<Expr> &ex) {
  if (!n) return n;
  if (n->kind == ExprNode::Input) {
    if (n->input >= (int)ex.size()) throw ArgumentError("expression input out of range");
    const Expr &r = ex[(size_t)n->input];
    if (compute_dtype(n->storage) != r.dtype()) return cast_to(r, compute_dtype(n->storage)).node;
    return r.node;
  }
  if (n->kind == ExprNode::Const || n->kind == ExprNode::Index) return n;
  ExprNode c = *n;
  c.a = subst(n->a, ex);
  c.b = subst(n->b, ex);
  c.c = subst(n->c, ex);
  return mk(c);
}

Expr Expr::substitute(const std::vector<Expr> &exprs) const { return Expr{subst(node, exprs)}; }

static int arity_of(const std::shared_ptr<const ExprNode> &n) {
  if (!n) return 0;
  if (n->kind == ExprNode::Input) return n->input + 1;
  return std::max(arity_of(n->a), std::max(arity_of(n->b), arity_of(n->c)));
}
int Expr::arity() const { return arity_of(node); }

static std::string to_s_node(const std::shared_ptr<const ExprNode> &n) {
  if (n->kind == ExprNode::Input) return "x" + std::to_string(n->input);
  if (n->kind == ExprNode::Const) return n->value.to_s();
  if (n->kind == ExprNode::Index) return "idx";
  std::string s = "(op" + std::to_string(n->op) + " " + to_s_node(n->a);
  if (n->b) s += " " + to_s_node(n->b);
  if (n->c) s += " " + to_s_node(n->c);
  return s + ")";
}
std::string Expr::to_s() const { return to_s_node(node); }

namespace {
struct Lowering {
  ph_prog prog;
  bool used[PH_NREG] = {false};
  int alloc() {
    for (int r = 0; r < PH_NREG; ++r)
      if (!used[r]) {
        used[r] = true;
        return r;
      }
    throw ArgumentError("element expression needs more than 8 registers");
  }
  void release(int r) { used[r] = false; }
  void emit(const ph_insn &in) {
    if (prog.n_insn >= PH_MAX_INSN) throw ArgumentError("element expression too long");
    prog.insn[prog.n_insn++] = in;
  }
  static int need(const std::shared_ptr<const ExprNode> &n) {
    if (n->kind != ExprNode::Op) return 1;
    int na = need(n->a);
    if (n->nargs == 1) return na;
    int nb = need(n->b);
    int nc = n->nargs == 3 ? need(n->c) : 0;
    int m = std::max(na, std::max(nb, nc));
    return (na == nb) ? m + 1 : m;
  }
  int lower(const std::shared_ptr<const ExprNode> &n) {
    ph_insn in;
    std::memset(&in, 0, sizeof(in));
    if (n->kind == ExprNode::Input) {
      in.op = PH_OP_LOAD;
      in.dt = (uint8_t)compute_dtype(n->storage);
      in.a = (uint8_t)n->input;
      in.dst = (uint8_t)alloc();
      emit(in);
      return in.dst;
    }
    if (n->kind == ExprNode::Index) {
      in.op = PH_OP_INDEX;
      in.dt = (uint8_t)DType::I32;
      in.dst = (uint8_t)alloc();
      emit(in);
      return in.dst;
    }
    if (n->kind == ExprNode::Const) {
      in.op = PH_OP_CONST;
      in.dt = (uint8_t)n->dt;
      if (n->dt == DType::I32 || n->dt == DType::Bool) in.imm.i = n->value.i;
      else in.imm.f = n->value.f;
      in.dst = (uint8_t)alloc();
      emit(in);
      return in.dst;
    }
    int ra = -1, rb = -1, rc = -1;
    if (n->nargs == 1) {
      ra = lower(n->a);
    } else {
      if (need(n->b) > need(n->a)) {
        rb = lower(n->b);
        ra = lower(n->a);
      } else {
        ra = lower(n->a);
        rb = lower(n->b);
      }
      if (n->nargs == 3) rc = lower(n->c);
    }
    in.op = (uint8_t)n->op;
    in.dt = (uint8_t)(n->op == PH_OP_CAST ? n->dt : n->opdt);
    if (n->op == PH_OP_POWI) in.dt = (uint8_t)n->dt;
    in.a = (uint8_t)ra;
    in.b = (uint8_t)(rb < 0 ? 0 : rb);
    in.c = (uint8_t)(rc < 0 ? 0 : rc);
    if (n->op == PH_OP_CAST) in.c = (uint8_t)n->opdt;
    release(ra);
    if (rb >= 0) release(rb);
    if (rc >= 0) release(rc);
    in.dst = (uint8_t)alloc();
    emit(in);
    return in.dst;
  }
};
}  // namespace

ph_prog Expr::compile(int n_in) const {
  Lowering L;
  std::memset(&L.prog, 0, sizeof(L.prog));
  int r = L.lower(node);
  L.prog.n_in = n_in;
  L.prog.out_reg = r;
  L.prog.out_dtype = (int32_t)node->dt;
  return L.prog;
}

// ===========================================================================
// Storage / descriptors
// ===========================================================================
// Device buffers are recycled through a size-keyed pool: NArrays created and
// dropped in a loop (gather outputs, fresh fills) reuse memory instead of
// paying hipMalloc/hipFree (which synchronises) every time.  All device work
// runs on the library's single stream, so a recycled buffer is only touched
// by work ordered after everything that used it before.
namespace {
struct Pool {
  std::multimap<size_t, void *> free_list;
  size_t cached = 0;
  static constexpr size_t kCap = size_t(16) << 30;
};
Pool &pool() {
  static Pool *p = new Pool();  // never destroyed: buffers die with the process
  return *p;
}
size_t round_bytes(size_t b) {
  if (b < 512) return 512;
  size_t unit = b < (size_t(1) << 20) ? 512 : (size_t(1) << 20);
  return (b + unit - 1) / unit * unit;
}
void *pool_alloc(size_t bytes) {
  Pool &p = pool();
  auto it = p.free_list.find(bytes);
  if (it != p.free_list.end()) {
    void *ptr = it->second;
    p.free_list.erase(it);
    p.cached -= bytes;
    return ptr;
  }
  void *ptr = nullptr;
  int rc = ph_alloc(&ptr, bytes);
  if (rc == PH_E_OOM && !p.free_list.empty()) {  // give cached memory back and retry
    for (auto &kv : p.free_list) ph_free(kv.second);
    p.free_list.clear();
    p.cached = 0;
    rc = ph_alloc(&ptr, bytes);
  }
  check(rc, "NArray allocation");
  return ptr;
}
void pool_release(void *ptr, size_t bytes) {
  Pool &p = pool();
  if (p.cached + bytes > Pool::kCap) {
    ph_free(ptr);
    return;
  }
  p.free_list.emplace(bytes, ptr);
  p.cached += bytes;
}
}  // namespace

Storage::~Storage() {
  if (dptr) pool_release(dptr, round_bytes(bytes));
}

static ph_view_desc contiguous_desc(DType dt, const Shape &shape, int64_t extent) {
  ph_view_desc d;
  std::memset(&d, 0, sizeof(d));
  d.dtype = (int32_t)dt;
  d.extent = extent;
  d.view.rank = (int32_t)shape.size();
  int64_t st = 1;
  for (int i = (int)shape.size() - 1; i >= 0; --i) {
    d.view.shape[i] = shape[(size_t)i];
    d.view.stride[i] = st;
    st *= shape[(size_t)i];
  }
  return d;
}

static MirrorStats g_mirror;
MirrorStats mirror_stats() { return g_mirror; }

// Whether any element of the program can raise a flag (checked Int32
// arithmetic, floored division or modulo, a negative Int32 exponent).
static bool prog_may_raise(const ph_prog &p) {
  // registers holding an Int32 constant (a divisor known when the program
  // is built: // and % by one other than 0 and -1 cannot raise)
  bool is_const[256] = {false};
  int64_t cval[256] = {0};
  for (int k = 0; k < p.n_insn; ++k) {
    const ph_insn &in = p.insn[k];
    const bool konst = in.op == PH_OP_CONST && in.dt == PH_I32;
    if (in.dt == PH_I32) {
      switch (in.op) {
        case PH_OP_ADD: case PH_OP_SUB: case PH_OP_MUL: case PH_OP_IPOW: case PH_OP_NEG: case PH_OP_ABS:
          return true;
        case PH_OP_FLOORDIV: case PH_OP_MOD:
          if (!is_const[in.b] || cval[in.b] == 0 || cval[in.b] == -1) return true;
          break;
        default:
          break;
      }
    }
    is_const[in.dst] = konst;
    cval[in.dst] = konst ? in.imm.i : 0;
  }
  return false;
}

void Storage::run_pending_map() {
  std::shared_ptr<PendingMap> pm = pending_map;  // kept until the map has been issued
  if (!pm || elements == 0) {
    pending_map.reset();
    return;
  }
  std::vector<const void *> ptrs;
  std::vector<ph_view_desc> descs;
  for (const Operand &o : pm->inputs) {
    ptrs.push_back(o.storage->dev_read());
    descs.push_back(o.desc);
  }
  const ph_prog prog = pm->body.compile((int)pm->inputs.size());
  int32_t flags = 0;
  check(ph_map(&prog, (int32_t)pm->inputs.size(), ptrs.data(), descs.data(), dptr, &pm->out, &flags),
        "map (pending operator result)");
  pending_map.reset();
}

void Storage::before_write() {
  if (readers.empty()) return;
  std::vector<std::weak_ptr<Storage>> rs;
  rs.swap(readers);
  for (const auto &w : rs)
    if (auto r = w.lock())
      if (r->pending_map) r->run_pending_map();
}

const void *Storage::dev_read() {
  if (pending_map) run_pending_map();
  if (pending_fill) {
    Scalar v = *pending_fill;
    pending_fill.reset();
    if (elements > 0) {
      ph_view_desc d = contiguous_desc(dtype, {elements}, elements);
      unsigned char bits[8];
      scalar_bits(v, dtype, bits);
      check(ph_fill(dptr, &d, bits), "NArray.fill");
    }
  }
  if (dirty_) {
    // write back each run of dirty pages with one copy
    const size_t np = page_.size();
    for (size_t p = 0; p < np;) {
      if (page_[p] != 2) {
        ++p;
        continue;
      }
      size_t q = p;
      while (q < np && page_[q] == 2) page_[q++] = 1;
      const size_t off = p * kPage, len = std::min(bytes, q * kPage) - off;
      check(ph_h2d((char *)dptr + off, host_.data() + off, len), "host mirror write-back");
      ++g_mirror.runs_written;
      g_mirror.bytes_written += (int64_t)len;
      p = q;
    }
    dirty_ = 0;
  }
  return dptr;
}

void Storage::forget_host() {
  if (dirty_) throw ArgumentError("host mirror: dirty pages would be lost");
  if (held_) {
    std::fill(page_.begin(), page_.end(), (uint8_t)0);
    held_ = 0;
  }
}

void *Storage::dev_write() {
  before_write();
  dev_read();
  if (held_) {
    std::fill(page_.begin(), page_.end(), (uint8_t)0);
    held_ = 0;
  }
  return dptr;
}

void Storage::ensure_pages(size_t off, size_t len) {
  if (off + len > bytes) throw IndexError("host mirror access out of range");
  if (pending_map) run_pending_map();
  if (host_.size() != bytes) {
    host_.assign(bytes, 0);
    page_.assign((bytes + kPage - 1) / kPage, (uint8_t)0);
    held_ = dirty_ = 0;
  }
  if (len == 0) return;
  for (size_t p = off / kPage; p <= (off + len - 1) / kPage; ++p) {
    if (page_[p]) continue;
    const size_t po = p * kPage, pl = std::min(bytes, po + kPage) - po;
    if (pending_fill) {
      // the device has not been filled yet: the page takes the fill value
      unsigned char bits[8];
      scalar_bits(*pending_fill, dtype, bits);
      const size_t es = dtype_size(dtype);
      for (size_t b = 0; b < pl; b += es) std::memcpy(host_.data() + po + b, bits, es);
    } else {
      check(ph_d2h(host_.data() + po, (const char *)dptr + po, pl), "host mirror fetch");
      ++g_mirror.pages_fetched;
      g_mirror.bytes_fetched += (int64_t)pl;
    }
    page_[p] = 1;
    ++held_;
  }
}

const unsigned char *Storage::host_read(size_t off, size_t len) {
  ensure_pages(off, len);
  return host_.data() + off;
}

unsigned char *Storage::host_write(size_t off, size_t len) {
  before_write();
  ensure_pages(off, len);
  if (len > 0)
    for (size_t p = off / kPage; p <= (off + len - 1) / kPage; ++p)
      if (page_[p] != 2) {
        page_[p] = 2;
        ++dirty_;
      }
  return host_.data() + off;
}

static std::shared_ptr<Storage> new_storage(DType dt, int64_t elements) {
  init();
  auto s = std::make_shared<Storage>();
  s->dtype = dt;
  s->elements = elements;
  s->bytes = (size_t)elements * dtype_size(dt);
  s->dptr = pool_alloc(round_bytes(s->bytes));
  return s;
}

static void check_shape(const Shape &shape) {
  for (int64_t d : shape)
    if (d < 0) throw DimensionError("Cannot create NArray: One or more of the provided dimensions was negative.");
  if (shape.size() > PH_MAX_RANK) throw DimensionError("at most 8 dimensions are supported on the device");
}

// host evaluation of a descriptor at one coordinate (element access)
static int64_t desc_index(const ph_view_desc &d, const Coord &c) {
  int64_t idx = d.view.offset;
  for (int i = 0; i < d.view.rank; ++i) idx += c[(size_t)i] * d.view.stride[i];
  for (int s = 0; s < d.nstage; ++s) {
    const ph_affine &st = d.stage[s];
    int64_t out = st.offset;
    for (int j = st.rank - 1; j >= 0; --j) {
      int64_t q = idx / st.shape[j];
      out += (idx - q * st.shape[j]) * st.stride[j];
      idx = q;
    }
    idx = out;
  }
  return idx;
}

static Scalar fetch_one(const std::shared_ptr<Storage> &st, int64_t idx) {
  const size_t es = dtype_size(st->dtype);
  return scalar_from_bits(st->dtype, st->host_read((size_t)idx * es, es));
}

static void store_one(const std::shared_ptr<Storage> &st, int64_t idx, const Scalar &v) {
  const size_t es = dtype_size(st->dtype);
  unsigned char bits[8];
  scalar_bits(v, st->dtype, bits);
  std::memcpy(st->host_write((size_t)idx * es, es), bits, es);
}

// --- descriptor composition (a new transform applied to view coordinates
// first, then the existing descriptor) --------------------------------------
static void compose_region(ph_view_desc &d, const IndexRegion &r) {
  ph_affine &v = d.view;
  int64_t off = v.offset;
  ph_affine nv;
  std::memset(&nv, 0, sizeof(nv));
  int k = 0;
  for (int i = 0; i < v.rank; ++i) {
    off += r.first[(size_t)i] * v.stride[i];
    if (!(r.drop && r.degeneracy[(size_t)i])) {
      nv.shape[k] = r.proper_shape[(size_t)i];
      nv.stride[k] = r.step[(size_t)i] * v.stride[i];
      ++k;
    }
  }
  if (k == 0) {  // every axis dropped: one axis of 1 (or 0) elements
    nv.shape[0] = r.reduced_shape[0];
    nv.stride[0] = 0;
    k = 1;
  }
  nv.rank = k;
  nv.offset = off;
  // an empty axis (step 0) addresses nothing; keep offsets in range
  for (int i = 0; i < k; ++i)
    if (nv.shape[i] == 0) nv.offset = 0;
  v = nv;
}

static void compose_permute(ph_view_desc &d, const std::vector<int64_t> &pattern) {
  ph_affine nv = d.view;
  for (size_t i = 0; i < pattern.size(); ++i) {
    nv.shape[i] = d.view.shape[pattern[i]];
    nv.stride[i] = d.view.stride[pattern[i]];
  }
  d.view = nv;
}

static void compose_reverse(ph_view_desc &d) {
  for (int i = 0; i < d.view.rank; ++i) {
    if (d.view.shape[i] > 0) d.view.offset += (d.view.shape[i] - 1) * d.view.stride[i];
    d.view.stride[i] = -d.view.stride[i];
  }
}

static bool view_is_rowmajor(const ph_affine &v) {
  int64_t st = 1;
  for (int i = v.rank - 1; i >= 0; --i) {
    if (v.shape[i] != 1 && v.stride[i] != st) return false;
    st *= v.shape[i];
  }
  return true;
}

// Adds a reshape/tile stage; false (d unchanged) when the descriptor has no
// room for another, and the view then keeps the transform for later.
static bool push_stage(ph_view_desc &d, const ph_affine &stage_map, const ph_affine &new_view) {
  if (d.nstage >= PH_MAX_STAGES) return false;
  for (int s = d.nstage; s > 0; --s) d.stage[s] = d.stage[s - 1];
  d.stage[0] = stage_map;
  d.nstage += 1;
  d.view = new_view;
  return true;
}

static bool compose_reshape(ph_view_desc &d, const Shape &new_shape) {
  ph_affine nv;
  std::memset(&nv, 0, sizeof(nv));
  nv.rank = (int32_t)new_shape.size();
  int64_t st = 1;
  for (int i = nv.rank - 1; i >= 0; --i) {
    nv.shape[i] = new_shape[(size_t)i];
    nv.stride[i] = st;
    st *= new_shape[(size_t)i];
  }
  if (view_is_rowmajor(d.view)) {  // linear index maps straight through
    nv.offset = d.view.offset;
    d.view = nv;
    return true;
  }
  return push_stage(d, d.view, nv);
}

static bool compose_tile(ph_view_desc &d, const Shape &src_shape, const std::vector<int64_t> &counts) {
  bool need_mod = false;
  for (size_t i = 0; i < src_shape.size(); ++i)
    if (src_shape[i] != 1 && counts[i] != 1) need_mod = true;
  ph_affine nv = d.view;
  for (size_t i = 0; i < src_shape.size(); ++i) {
    nv.shape[i] = src_shape[i] * counts[i];
    if (src_shape[i] == 1) nv.stride[i] = 0;  // broadcast
  }
  if (!need_mod) {
    d.view = nv;
    return true;
  }
  if (2 * src_shape.size() > PH_MAX_RANK) return false;
  // linearise the tiled coordinate, delinearise as [k0, s0, k1, s1, ...]
  ph_affine stage;
  std::memset(&stage, 0, sizeof(stage));
  stage.rank = (int32_t)(2 * src_shape.size());
  stage.offset = d.view.offset;
  for (size_t i = 0; i < src_shape.size(); ++i) {
    stage.shape[2 * i] = counts[i];
    stage.shape[2 * i + 1] = src_shape[i];
    stage.stride[2 * i] = 0;
    stage.stride[2 * i + 1] = d.view.stride[i];
  }
  ph_affine lin;
  std::memset(&lin, 0, sizeof(lin));
  lin.rank = (int32_t)src_shape.size();
  int64_t st = 1;
  for (int i = lin.rank - 1; i >= 0; --i) {
    lin.shape[i] = nv.shape[i];
    lin.stride[i] = st;
    st *= nv.shape[i];
  }
  return push_stage(d, stage, lin);
}

// ===========================================================================
// NArray
// ===========================================================================
NArray NArray::empty(const Shape &shape, DType dtype) {
  check_shape(shape);
  NArray a;
  a.shape_ = shape;
  a.storage = new_storage(dtype, ShapeUtil::shape_to_size(shape));
  return a;
}

NArray NArray::fill(const Shape &shape, const Scalar &value, std::optional<DType> dtype) {
  NArray a = empty(shape, dtype ? *dtype : value.dt);
  a.storage->pending_fill = value;
  return a;
}

NArray NArray::of_buffer(const Shape &shape, DType dtype, const void *host, size_t bytes) {
  check_shape(shape);
  const int64_t n = ShapeUtil::shape_to_size(shape);
  if ((size_t)n * dtype_size(dtype) != bytes)
    throw ShapeError("Cannot create NArray: Given shape does not match number of elements in buffer.");
  NArray a = empty(shape, dtype);
  check(ph_h2d(a.storage->dev_write(), host, bytes), "NArray.of_buffer");
  return a;
}

NArray NArray::random(const Shape &shape, DType dtype, uint64_t seed, uint64_t offset, int dist, double a,
                      double b) {
  NArray r = empty(shape, dtype);
  check(ph_fill_random(r.storage->dev_write(), (int32_t)dtype, r.size(), seed, offset, dist, a, b), "random fill");
  return r;
}

void *NArray::device_ptr() const { return storage->dev_write(); }

ph_view_desc NArray::desc() const {
  return contiguous_desc(dtype(), shape_, std::max<int64_t>(storage->elements, 1));
}

Operand NArray::operand() const {
  Operand o;
  o.storage = storage;
  o.desc = desc();
  o.shape = shape_;
  return o;
}

DType Operand::value_dtype() const { return body ? body->dtype() : compute_dtype((DType)desc.dtype); }

void NArray::to_host(void *out, size_t bytes) const {
  if (bytes != (size_t)size() * dtype_size(dtype())) throw ArgumentError("to_host: wrong byte count");
  check(ph_d2h(out, storage->dev_read(), bytes), "NArray#to_host");
}

Scalar NArray::get(const Coord &coord) const {
  Coord c = CoordUtil::canonicalize_coord(coord, shape_);
  return fetch_one(storage, desc_index(desc(), c));
}

void NArray::set_element(const Coord &coord, const Scalar &v) {
  Coord c = CoordUtil::canonicalize_coord(coord, shape_);
  store_one(storage, desc_index(desc(), c), v);
}

static NArray map_core(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn,
                       const NArray *out_in);
static bool map_into(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn,
                     const std::shared_ptr<Storage> &dst, const ph_view_desc &od);

// gather of an arbitrary operand (view / ProcView) into a new NArray
static NArray materialize_operand(const Operand &o) {
  if (o.body) {
    // ProcView: a map of its block
    return map_with({o}, [](const std::vector<Expr> &x) { return x[0]; });
  }
  NArray out = NArray::empty(o.shape, (DType)o.desc.dtype);
  if (out.size() == 0) return out;
  const void *src = o.storage->dev_read();
  ph_view_desc dd = out.desc();
  check(ph_strided_copy(out.storage->dev_write(), &dd, src, &o.desc, 0), "gather");
  return out;
}

static void apply_transform(View &v, const Transform &t);

// Buffer indices of the elements of a view descriptor, lexicographic, as
// int64 values in an array of 8-byte elements.
static NArray index_of_desc(const ph_view_desc &d) {
  NArray r = NArray::empty(Shape(d.view.shape, d.view.shape + d.view.rank), DType::F64);
  if (r.size() > 0) check(ph_view_index((int64_t *)r.storage->dev_write(), &d), "view index");
  return r;
}

// The same for a chain past one descriptor (then restricted to `region`):
// the index array of its descriptor, gathered through the rest of the chain
// (8-byte copies move the indices unchanged).
static NArray index_of_chain(const DeepChain &c, const IndexRegion *region = nullptr) {
  NArray mid = index_of_desc(c.desc);
  View v = mid.view();
  for (size_t i = c.tail.size(); i-- > 0;) apply_transform(v, c.tail[i]);
  if (region) v.restrict_to(*region);
  return v.to_narr();
}

// dst's elements (through descriptor dd, or at the buffer indices in `wi`)
// <- src, element k of one from element k of the other, both lexicographic,
// with the source read element by element in the reference's loop order
// (src/n_array.cr:484-492 pulls `src.each` as it writes), so it sees the
// writes made before it.  Used where the source reads dst's own buffer and
// either side is a view past one descriptor.  False (nothing written) for a
// ProcView or conversion into BFloat16 storage, which only a gathered copy
// can round.
static bool write_in_order(const std::shared_ptr<Storage> &dst, const ph_view_desc &dd, const NArray *wi,
                           const Operand &src) {
  const std::shared_ptr<Storage> &ss = src.deep ? src.deep->storage : src.storage;
  const DType sdt = ss->dtype, ddt = dst->dtype;
  const bool copy = !src.body && dtype_size(sdt) == dtype_size(ddt) && compute_dtype(sdt) == compute_dtype(ddt);
  ph_prog prog{};
  if (!copy) {
    Expr e = Expr::input(0, sdt);
    if (src.body) e = src.body->substitute({e});
    if (e.dtype() != compute_dtype(ddt))
      throw ArgumentError(std::string("cannot store ") + dtype_name(e.dtype()) + " values into a " + dtype_name(ddt) +
                          " array");
    if (ddt == DType::BF16) return false;
    prog = e.compile(1);
  }
  NArray ri = src.deep ? index_of_chain(*src.deep) : index_of_desc(src.desc);
  NArray wtmp;
  if (!wi && dd.nstage != 0) {
    wtmp = index_of_desc(dd);
    wi = &wtmp;
  }
  // indexed operands: their descriptors give dtype, extent and shape
  ph_view_desc sd = src.deep ? src.deep->desc : src.desc;
  sd.dtype = (int32_t)sdt;
  sd.view.rank = dd.view.rank;
  for (int i = 0; i < dd.view.rank; ++i) sd.view.shape[i] = dd.view.shape[i];
  const void *ip = ss->dev_read();
  const int64_t *rip = (const int64_t *)ri.storage->dev_read();
  const int64_t *wip = wi ? (const int64_t *)wi->storage->dev_read() : nullptr;
  void *op = dst->dev_write();
  int32_t flags = 0;
  check(ph_map_indexed(copy ? nullptr : &prog, 1, &ip, &sd, &rip, op, &dd, wip, &flags), "ordered write");
  check_flags(flags, "elementwise");
  return true;
}

NArray NArray::unsafe_fetch_chunk(const IndexRegion &region) const {
  View v = view();
  v.restrict_to(region);
  return v.to_narr();
}

NArray NArray::get_chunk(const IndexRegion &region) const {
  if (region.proper_dimensions() != dimensions())
    throw DimensionError("'region' was " + std::to_string(region.proper_dimensions()) +
                         "-dimensional, but this MultiIndexable is " + std::to_string(dimensions()) + "-dimensional.");
  if (!region.fits_in(shape_))
    throw ShapeError("'region' (" + region.to_s() + ") cannot fit into a MultiIndexable with shape " + shape_to_s(shape_) + ".");
  return unsafe_fetch_chunk(region);
}

NArray NArray::get_chunk(const RegionLiteral &lit, bool drop) const {
  return unsafe_fetch_chunk(IndexRegion::from_literal(lit, shape_, drop));
}

NArray NArray::get_available(const RegionLiteral &lit, bool drop) const {
  return unsafe_fetch_chunk(IndexRegion::trimmed(lit, &shape_, drop, shape_));
}

View NArray::view() const {
  View v;
  v.storage = storage;
  v.src_shape = shape_;
  v.shape_ = shape_;
  v.desc_ = desc();
  return v;
}

View NArray::view(const RegionLiteral &lit) const {
  View v = view();
  v.restrict_to(lit);
  return v;
}

MutableView NArray::mutable_view() const {
  MutableView m;
  static_cast<View &>(m) = view();
  return m;
}

MutableView NArray::mutable_view(const RegionLiteral &lit) const {
  MutableView m = mutable_view();
  m.restrict_to(lit);
  return m;
}

ProcView NArray::process(const Expr &body) const { return view().process(body); }

View NArray::tile(const std::vector<int64_t> &counts) const {
  View v = view();
  v.tile_(counts);
  return v;
}

NArray NArray::reshape(const Shape &s) const {
  check_shape(s);
  if (ShapeUtil::shape_to_size(s) != size())
    throw ShapeError("Cannot create NArray: Given shape does not match number of elements in buffer.");
  NArray r = *this;
  r.shape_ = s;
  return r;
}

NArray NArray::flatten() const { return reshape({size()}); }

NArray NArray::permute(const std::vector<int64_t> *order) const { return view().permute(order).to_narr(); }
NArray NArray::reverse() const { return view().reverse().to_narr(); }
NArray NArray::clone() const { return view().to_narr(); }

static void scatter_into(const std::shared_ptr<Storage> &dst, const ph_view_desc &dd, const Shape &buf_shape,
                         const Operand &src_in);

// compatible?(*others, axis) (n_array.cr:666-673): every dimension matches,
// except at index `axis` (as given: a negative axis exempts no dimension).
// Inputs of another rank are rejected with the same error.
static void check_concat(const std::vector<Operand> &narrs, int64_t axis) {
  const Shape &s0 = narrs[0].shape;
  const std::string msg = "Cannot concatenate these arrays along axis " + std::to_string(axis) +
                          ": shapes do not match";
  for (const Operand &n : narrs) {
    if (n.shape.size() != s0.size()) throw DimensionError(msg);
    for (size_t idx = 0; idx < s0.size(); ++idx)
      if (n.shape[idx] != s0[idx] && (int64_t)idx != axis) throw DimensionError(msg);
  }
}

// NArray.concatenate_to_slice (n_array.cr:727-751): the output takes the
// common shape with `axis` summed; input i fills the slab of `axis`
// coordinates [sum of the earlier inputs' extents, + its extent).  Each slab
// is one strided scatter, so a chunk of consecutive output elements is
// written from consecutive input elements, in the reference's order.
NArray NArray::concatenate(const std::vector<Operand> &narrs, int64_t axis) {
  if (narrs.empty()) throw ArgumentError("concatenate needs at least one MultiIndexable");
  check_concat(narrs, axis);
  const DType vt = narrs[0].value_dtype();
  for (const Operand &n : narrs)
    if (n.value_dtype() != vt) throw ArgumentError("concatenate: every MultiIndexable must have the same element type");
  const int64_t rank = (int64_t)narrs[0].shape.size();
  const int64_t ax = axis < 0 ? axis + rank : axis;  // Crystal's negative indexing of the shape
  if (ax < 0 || ax >= rank) throw IndexError("Index out of bounds");
  Shape out_shape = narrs[0].shape;
  out_shape[(size_t)ax] = 0;
  for (const Operand &n : narrs) out_shape[(size_t)ax] += n.shape[(size_t)ax];
  DType st = vt;
  if (vt == compute_dtype((DType)narrs[0].desc.dtype) && !narrs[0].body) st = (DType)narrs[0].desc.dtype;
  NArray out = NArray::empty(out_shape, st);
  int64_t off = 0;
  for (const Operand &n : narrs) {
    const int64_t len = n.shape[(size_t)ax];
    if (len > 0 && out.size() > 0) {
      View v = out.view();
      IndexRegion r = IndexRegion::cover(out_shape);
      r.first[(size_t)ax] = off;
      r.last[(size_t)ax] = off + len - 1;
      r.proper_shape[(size_t)ax] = len;
      r.reduced_shape = r.proper_shape;
      v.restrict_to(r);
      scatter_into(out.storage, v.desc_, out_shape, n);
    }
    off += len;
  }
  return out;
}

NArray NArray::concatenate_with(const std::vector<Operand> &others, int64_t axis) const {
  std::vector<Operand> all{operand()};
  all.insert(all.end(), others.begin(), others.end());
  return NArray::concatenate(all, axis);
}

NArray &NArray::concatenate_inplace(const std::vector<Operand> &others, int64_t axis) {
  NArray r = concatenate_with(others, axis);
  storage = r.storage;
  shape_ = r.shape_;
  return *this;
}

// push(*others, axis = 0) (n_array.cr:688-710): in-place along axis 0
NArray &NArray::push(const std::vector<Operand> &others) { return concatenate_inplace(others, 0); }

std::vector<NArray> NArray::slices(int axis) const {
  if (axis < 0 || axis >= dimensions()) throw IndexError("axis out of range");
  Shape chunk = shape_;
  chunk[(size_t)axis] = 1;
  std::vector<bool> degen((size_t)dimensions(), false);
  degen[(size_t)axis] = true;
  std::vector<NArray> out;
  for (int64_t j = 0; j < shape_[(size_t)axis]; ++j) {
    IndexRegion r = IndexRegion::cover(chunk, true, &degen);
    std::vector<int64_t> off((size_t)dimensions(), 0);
    off[(size_t)axis] = j;
    r.translate(off);
    out.push_back(unsafe_fetch_chunk(r));
  }
  return out;
}

// ---------------------------------------------------------------------------
// scatter
// ---------------------------------------------------------------------------

// Reshape `src` to the destination's logical shape when the two are
// compatible up to trailing ones (ShapeUtil.compatible_shapes?).
static std::optional<ph_view_desc> fit_src_desc(const ph_view_desc &src, const Shape &dst_shape) {
  ph_view_desc d = src;
  if (d.nstage > 0) {
    // re-express the (equal-size) source in the destination shape: reshape it
    if (!compose_reshape(d, dst_shape)) return std::nullopt;
    return d;
  }
  // same leading extents; drop or add trailing unit axes
  ph_affine nv;
  std::memset(&nv, 0, sizeof(nv));
  nv.rank = (int32_t)dst_shape.size();
  nv.offset = src.view.offset;
  for (size_t i = 0; i < dst_shape.size(); ++i) {
    nv.shape[i] = dst_shape[i];
    nv.stride[i] = (int)i < src.view.rank ? src.view.stride[i] : 0;
  }
  d.view = nv;
  return d;
}

// Whether the destination descriptor is a strided box of the row-major
// buffer (a region, possibly permuted/reversed): per buffer axis a, first,
// step, count and which view axis walks it (-1 fixed).
struct BoxForm {
  std::vector<int64_t> first, step, count;
  std::vector<int32_t> axis;
};
static std::optional<BoxForm> box_form(const ph_view_desc &d, const Shape &buf_shape) {
  if (d.nstage != 0) return std::nullopt;
  const size_t R = buf_shape.size();
  if (R == 0) return std::nullopt;
  std::vector<int64_t> rm(R);
  int64_t st = 1;
  for (int i = (int)R - 1; i >= 0; --i) {
    rm[(size_t)i] = st;
    st *= buf_shape[(size_t)i];
  }
  BoxForm f;
  f.first.assign(R, 0);
  f.step.assign(R, 1);
  f.count.assign(R, 1);
  f.axis.assign(R, -1);
  int64_t off = d.view.offset;
  for (size_t a = 0; a < R; ++a) {
    f.first[a] = off / rm[a];
    off -= f.first[a] * rm[a];
  }
  for (int i = 0; i < d.view.rank; ++i) {
    const int64_t n = d.view.shape[i], s = d.view.stride[i];
    if (n == 1) continue;
    if (n == 0 || s == 0) return std::nullopt;
    bool found = false;
    for (size_t a = 0; a < R && !found; ++a) {
      if (f.axis[a] >= 0 || s % rm[a] != 0) continue;
      const int64_t step = s / rm[a];
      const int64_t last = f.first[a] + (n - 1) * step;
      if (last < 0 || last >= buf_shape[a]) continue;
      f.axis[a] = i;
      f.step[a] = step;
      f.count[a] = n;
      found = true;
    }
    if (!found) return std::nullopt;
  }
  // unit view axes map onto fixed buffer axes; give them a view axis when a
  // buffer axis is free so the source's unit axes are accounted for
  for (int i = 0; i < d.view.rank; ++i) {
    if (d.view.shape[i] != 1) continue;
    for (size_t a = 0; a < R; ++a)
      if (f.axis[a] < 0) {
        f.axis[a] = i;
        f.count[a] = 1;
        f.step[a] = 1;
        break;
      }
  }
  // verify: the box must reproduce the descriptor
  int64_t o = 0;
  for (size_t a = 0; a < R; ++a) o += f.first[a] * rm[a];
  if (o != d.view.offset) return std::nullopt;
  for (int i = 0; i < d.view.rank; ++i) {
    bool mapped = false;
    for (size_t a = 0; a < R; ++a)
      if (f.axis[a] == i) {
        mapped = true;
        if (d.view.shape[i] != 1 && f.step[a] * rm[a] != d.view.stride[i]) return std::nullopt;
      }
    if (!mapped) return std::nullopt;
  }
  return f;
}

// dst (storage + view descriptor over buf_shape) <- src, elementwise over the
// destination's logical shape.
static void scatter_into(const std::shared_ptr<Storage> &dst, const ph_view_desc &dd, const Shape &buf_shape,
                         const Operand &src_in) {
  int64_t n = 1;
  for (int i = 0; i < dd.view.rank; ++i) n *= dd.view.shape[i];
  if (n == 0) return;
  Shape dshape(dd.view.shape, dd.view.shape + dd.view.rank);
  if (src_in.deep && src_in.deep->storage == dst && write_in_order(dst, dd, nullptr, src_in)) return;
  Operand src = src_in;
  std::optional<ph_view_desc> fd = fit_src_desc(src.desc, dshape);
  if (!fd) {  // the source reshaped to dst's shape needs one stage more: gather it first
    NArray tmp = materialize_operand(src);
    src = tmp.operand();
    fd = fit_src_desc(src.desc, dshape);
  }
  if (src.body || compute_dtype((DType)src.desc.dtype) != compute_dtype((DType)dd.dtype) ||
      dtype_size((DType)src.desc.dtype) != dtype_size((DType)dd.dtype)) {
    // a ProcView (or a different element type): evaluated element by element
    // into the destination, in the reference's order where it reads it
    DType target = (DType)dd.dtype;
    if (src.value_dtype() != compute_dtype(target))
      throw ArgumentError(std::string("cannot store ") + dtype_name(src.value_dtype()) + " values into a " +
                          dtype_name(target) + " array");
    Operand fitted = src;
    fitted.desc = *fd;
    fitted.shape = dshape;
    if (map_into({fitted}, [](const std::vector<Expr> &x) { return x[0]; }, dst, dd)) return;
    NArray tmp = materialize_operand(src);  // BF16 storage: rounded on the way
    src = tmp.operand();
    fd = fit_src_desc(src.desc, dshape);
  }
  const ph_view_desc sd = *fd;
  const void *sp = src.storage->dev_read();
  if (dst->pending_fill && src.storage != dst && dst->host_clean()) {
    if (auto f = box_form(dd, buf_shape)) {
      // the pending fill and the scatter in one pass: every element of dst
      // is written, so the (clean) host mirror is dropped
      unsigned char bits[8];
      scalar_bits(*dst->pending_fill, dst->dtype, bits);
      dst->before_write();
      check(ph_fill_scatter(dst->dptr, (int32_t)dst->dtype, buf_shape.data(), (int32_t)buf_shape.size(),
                            f->first.data(), f->step.data(), f->count.data(), f->axis.data(), sp, &sd, bits),
            "fill + scatter");
      dst->pending_fill.reset();
      dst->forget_host();
      return;
    }
  }
  check(ph_strided_copy(dst->dev_write(), &dd, sp, &sd, 0), "scatter");
}

void NArray::unsafe_set_chunk(const IndexRegion &region, const Operand &src) {
  View v = view();
  v.restrict_to(region);
  scatter_into(storage, v.desc_, shape_, src);
}

void NArray::unsafe_set_chunk(const IndexRegion &region, const Scalar &value) {
  View v = view();
  v.restrict_to(region);
  if (v.size() == 0) return;
  unsigned char bits[8];
  scalar_bits(value, dtype(), bits);
  check(ph_fill(storage->dev_write(), &v.desc_, bits), "set_chunk (scalar)");
}

void NArray::set_chunk(const RegionLiteral &lit, const Operand &src) {
  IndexRegion r = IndexRegion::from_literal(lit, shape_);
  if (!ShapeUtil::compatible_shapes(src.shape, r.shape()))
    throw ShapeError("Cannot substitute MultiIndexable: the given MultiIndexable has shape " + shape_to_s(src.shape) +
                     ", but region " + r.to_s() + " has shape " + shape_to_s(r.shape()) + ".");
  unsafe_set_chunk(r, src);
}

void NArray::set_chunk(const RegionLiteral &lit, const Scalar &value) {
  unsafe_set_chunk(IndexRegion::from_literal(lit, shape_), value);
}

void NArray::set_mask(const Operand &mask, const Scalar &value) {
  if (mask.shape != shape_) throw DimensionError("Cannot perform masking: mask shape does not match array shape.");
  if (mask.value_dtype() != DType::Bool) throw ArgumentError("mask must be Bool");
  // dst = mask ? value : dst   (one fused pass, in place)
  Operand self = operand();
  map_core({mask, self}, [&](const std::vector<Expr> &x) {
    return Expr::select(x[0], Expr::constant(value), x[1]);
  }, this);
}

void NArray::set_mask(const Operand &mask, const Operand &src) {
  if (mask.shape != shape_) throw DimensionError("Cannot perform masking: mask shape does not match array shape.");
  if (src.shape != shape_) throw ShapeError("Cannot perform masking: source shape " + shape_to_s(src.shape) +
                                            " does not match array shape " + shape_to_s(shape_) + ".");
  if (mask.value_dtype() != DType::Bool) throw ArgumentError("mask must be Bool");
  Operand self = operand();
  map_core({mask, src, self}, [&](const std::vector<Expr> &x) { return Expr::select(x[0], x[1], x[2]); }, this);
}

// ===========================================================================
// View / MutableView / ProcView
// ===========================================================================
// Folds transform `t` into the descriptor, or keeps it in the tail once the
// descriptor has no room for its stage.
void View::push(const Transform &t, const Shape &new_shape, const std::function<bool(ph_view_desc &)> &compose) {
  if (tail_ == 0) {
    ph_view_desc d = desc_;
    if (compose(d)) {
      desc_ = d;
    } else {
      mid_shape_ = shape_;
      tail_ = 1;
    }
  } else {
    ++tail_;
  }
  transforms.insert(transforms.begin(), t);
  shape_ = new_shape;
}

static void apply_transform(View &v, const Transform &t) {
  switch (t.kind) {
    case Transform::Region: v.restrict_to(t.region); break;
    case Transform::Reshape: v.reshape_(t.b); break;
    case Transform::Permute: v.permute_(&t.a); break;
    case Transform::Reverse: v.reverse_(); break;
    case Transform::Tile: v.tile_(t.b); break;
  }
}

// The tail's transforms, oldest first, on `on` (a view of the intermediate).
void View::replay_tail(View &on) const {
  for (size_t i = tail_; i-- > 0;) apply_transform(on, transforms[i]);
}

int64_t View::buffer_index(const Coord &c) const {
  if (tail_ == 0) return desc_index(desc_, c);
  const Coord b = apply(c);  // the whole chain, to a coordinate of the buffer's shape
  int64_t idx = 0, st = 1;
  for (size_t i = b.size(); i-- > 0;) {
    idx += b[i] * st;
    st *= src_shape[i];
  }
  return idx;
}

Operand View::operand() const {
  if (tail_ > 0) {
    // gathered now, when the view is used, so it reads the source as it is
    Operand mid;
    mid.storage = storage;
    mid.desc = desc_;
    mid.shape = mid_shape_;
    View v = materialize_operand(mid).view();
    replay_tail(v);
    Operand o = v.operand();
    auto c = std::make_shared<DeepChain>();
    c->storage = storage;
    c->desc = desc_;
    c->mid_shape = mid_shape_;
    c->tail.assign(transforms.begin(), transforms.begin() + (std::ptrdiff_t)tail_);
    o.deep = c;
    return o;
  }
  Operand o;
  o.storage = storage;
  o.desc = desc_;
  o.shape = shape_;
  return o;
}

View &View::restrict_to(const IndexRegion &region) {
  Transform t;
  t.kind = Transform::Region;
  t.region = region;
  push(t, region.shape(), [&](ph_view_desc &d) {
    compose_region(d, region);
    return true;
  });
  return *this;
}

View &View::restrict_to(const RegionLiteral &lit) {
  // ph-core prints "i'm getting called" here (src/view.cr:44); the debug
  // print is not reproduced.
  return restrict_to(IndexRegion::from_literal(lit, shape_));
}

View View::view(const RegionLiteral &lit) const {
  View v = *this;
  v.restrict_to(lit);
  return v;
}

View &View::reshape_(const Shape &ns) {
  if (ShapeUtil::shape_to_size(ns) != size())
    throw ShapeError("Cannot change shape from " + shape_to_s(shape_) + " (" + std::to_string(size()) +
                     " elements) to " + shape_to_s(ns) + " (" + std::to_string(ShapeUtil::shape_to_size(ns)) +
                     " elements) because reshape cannot add or remove elements.");
  if (ns.size() > PH_MAX_RANK) throw DimensionError("at most 8 dimensions");
  Transform t;
  t.kind = Transform::Reshape;
  t.a = shape_;
  t.b = ns;
  push(t, ns, [&](ph_view_desc &d) { return compose_reshape(d, ns); });
  return *this;
}

View &View::permute_(const std::vector<int64_t> *order) {
  std::vector<int64_t> pattern;
  if (order) {
    for (int64_t ax : *order)
      if (ax < 0 || ax >= dimensions())
        throw IndexError("Could not use pattern " + shape_to_s(*order) + " to permute: Axis " + std::to_string(ax) +
                         " is not present in a " + std::to_string(dimensions()) + "-dimensional MultiIndexable");
    pattern = *order;
    if ((int)pattern.size() != dimensions()) throw DimensionError("permutation pattern has the wrong number of axes");
    std::vector<bool> seen((size_t)dimensions(), false);
    for (int64_t ax : pattern) {
      if (seen[(size_t)ax]) throw ArgumentError("permutation pattern repeats axis " + std::to_string(ax));
      seen[(size_t)ax] = true;
    }
  } else {
    for (int i = 0; i < dimensions(); ++i) pattern.push_back(dimensions() - i - 1);
  }
  Shape ns(pattern.size());
  for (size_t i = 0; i < pattern.size(); ++i) ns[i] = shape_[(size_t)pattern[i]];
  Transform t;
  t.kind = Transform::Permute;
  t.a = pattern;
  push(t, ns, [&](ph_view_desc &d) {
    compose_permute(d, pattern);
    return true;
  });
  return *this;
}

View &View::reverse_() {
  Transform t;
  t.kind = Transform::Reverse;
  t.a = shape_;
  push(t, shape_, [](ph_view_desc &d) {
    compose_reverse(d);
    return true;
  });
  return *this;
}

View &View::tile_(const std::vector<int64_t> &counts) {
  if (counts.size() != shape_.size())
    throw DimensionError("tile counts " + shape_to_s(counts) + " do not match a " + std::to_string(dimensions()) +
                         "-dimensional MultiIndexable");
  for (int64_t c : counts)
    if (c < 0) throw ArgumentError("negative tile count");
  Shape ns(shape_.size());
  for (size_t i = 0; i < ns.size(); ++i) ns[i] = shape_[i] * counts[i];
  Transform t;
  t.kind = Transform::Tile;
  t.a = shape_;
  t.b = counts;
  const Shape cur = shape_;
  push(t, ns, [&](ph_view_desc &d) { return compose_tile(d, cur, counts); });
  return *this;
}

View View::reshape(const Shape &s) const {
  View v = *this;
  v.reshape_(s);
  return v;
}
View View::permute(const std::vector<int64_t> *order) const {
  View v = *this;
  v.permute_(order);
  return v;
}
View View::reverse() const {
  View v = *this;
  v.reverse_();
  return v;
}

NArray View::to_narr() const { return materialize_operand(operand()); }

ProcView View::process(const Expr &body) const {
  ProcView p;
  static_cast<View &>(p) = *this;
  if (body.arity() > 1) throw ArgumentError("a ProcView block takes one element");
  p.body = body;
  return p;
}

Coord View::apply(const Coord &coord) const {
  // ComposedTransform#apply (transforms.cr:77-91): newest first
  Coord c = coord;
  for (const Transform &t : transforms) {
    switch (t.kind) {
      case Transform::Region: c = t.region.local_to_absolute_unsafe(c); break;
      case Transform::Reshape: {
        Shape vs(t.b.size());
        int64_t st = 1;
        for (int i = (int)t.b.size() - 1; i >= 0; --i) {
          vs[(size_t)i] = st;
          st *= t.b[(size_t)i];
        }
        int64_t index = 0;
        for (size_t i = 0; i < c.size(); ++i) index += c[i] * vs[i];
        Coord o(t.a.size());
        for (int d = (int)t.a.size() - 1; d >= 0; --d) {
          o[(size_t)d] = floormod(index, t.a[(size_t)d]);
          index = floordiv(index, t.a[(size_t)d]);
        }
        c = o;
        break;
      }
      case Transform::Permute: {
        Coord o = c;
        for (size_t i = 0; i < t.a.size(); ++i) o[(size_t)t.a[i]] = c[i];
        c = o;
        break;
      }
      case Transform::Reverse:
        for (size_t i = 0; i < c.size(); ++i) c[i] = t.a[i] - 1 - c[i];
        break;
      case Transform::Tile:
        for (size_t i = 0; i < c.size(); ++i) c[i] = floormod(c[i], t.a[i]);
        break;
    }
  }
  return c;
}

Scalar View::get(const Coord &coord) const {
  Coord c = CoordUtil::canonicalize_coord(coord, shape_);
  return fetch_one(storage, buffer_index(c));
}

MutableView MutableView::mutable_view(const RegionLiteral &lit) const {
  MutableView m = *this;
  m.restrict_to(lit);
  return m;
}

// A write through a chain longer than one descriptor: the view the
// descriptor covers is gathered into a temporary, the tail of the chain is
// written through on the temporary, and the temporary goes back through the
// descriptor (its unwritten elements rewrite the values they hold).
void MutableView::deep_write(const std::function<void(MutableView &)> &fn) {
  Operand mid;
  mid.storage = storage;
  mid.desc = desc_;
  mid.shape = mid_shape_;
  NArray m = materialize_operand(mid);
  MutableView t = m.mutable_view();
  replay_tail(t);
  fn(t);
  scatter_into(storage, desc_, src_shape, m.operand());
}

void MutableView::unsafe_set_chunk(const IndexRegion &region, const Operand &src) {
  if (deep()) {
    if (src.storage == storage || (src.deep && src.deep->storage == storage)) {
      // the source reads the buffer written: written through the chain's
      // index array, element by element in the loop's order
      DeepChain wc;
      wc.storage = storage;
      wc.desc = desc_;
      wc.mid_shape = mid_shape_;
      wc.tail.assign(transforms.begin(), transforms.begin() + (std::ptrdiff_t)tail_);
      NArray wi = index_of_chain(wc, &region);
      ph_view_desc dd = contiguous_desc(storage->dtype, region.shape(), std::max<int64_t>(storage->elements, 1));
      if (wi.size() == 0 || write_in_order(storage, dd, &wi, src)) return;
    }
    deep_write([&](MutableView &t) { t.unsafe_set_chunk(region, src); });
    return;
  }
  View v = *this;
  v.restrict_to(region);
  scatter_into(storage, v.desc_, src_shape, src);
}

void MutableView::set_chunk(const RegionLiteral &lit, const Operand &src) {
  IndexRegion r = IndexRegion::from_literal(lit, shape_);
  if (!ShapeUtil::compatible_shapes(src.shape, r.shape()))
    throw ShapeError("Cannot substitute MultiIndexable: the given MultiIndexable has shape " + shape_to_s(src.shape) +
                     ", but region " + r.to_s() + " has shape " + shape_to_s(r.shape()) + ".");
  unsafe_set_chunk(r, src);
}

void MutableView::set_chunk(const RegionLiteral &lit, const Scalar &value) {
  if (deep()) {
    deep_write([&](MutableView &t) { t.set_chunk(lit, value); });
    return;
  }
  View v = *this;
  v.restrict_to(IndexRegion::from_literal(lit, shape_));
  if (v.size() == 0) return;
  unsigned char bits[8];
  scalar_bits(value, dtype(), bits);
  check(ph_fill(storage->dev_write(), &v.desc_, bits), "MutableView set_chunk (scalar)");
}

void MutableView::set_element(const Coord &coord, const Scalar &v) {
  Coord c = CoordUtil::canonicalize_coord(coord, shape_);
  store_one(storage, buffer_index(c), v);
}

Operand ProcView::operand() const {
  Operand o = View::operand();
  o.body = body;
  return o;
}

NArray ProcView::to_narr() const { return materialize_operand(operand()); }

ProcView ProcView::process(const Expr &next) const {
  ProcView p = *this;
  p.body = next.substitute({body});
  return p;
}

Scalar ProcView::get(const Coord &coord) const {
  // evaluate the block on the device for one element (a 1-element map)
  Coord c = CoordUtil::canonicalize_coord(coord, shape_);
  View one = *this;
  RegionLiteral lit;
  for (int64_t x : c) lit.push_back(RangeLiteral::of_index(x));
  one.restrict_to(IndexRegion::from_literal(lit, shape_));
  Operand o = one.operand();
  o.body = body;
  NArray r = materialize_operand(o);
  return r.get({0});
}

// ===========================================================================
// Elementwise
// ===========================================================================
// Fused map over several operands of one shape, into `out` (a contiguous
// NArray of the result type), or into a new NArray when out is null.
namespace {
// A fused map's kernel inputs (distinct storage and descriptor pairs) and
// its body over them.  An input read through a chain past one descriptor
// from the buffer the map writes has its index array in `index` (and its
// chain in `chain`); the others have empty entries there.
struct MapPlan {
  std::vector<Operand> inputs;
  std::vector<NArray> index;
  std::vector<const DeepChain *> chain;
  Expr body;
};

// Whether operand `o` reads the whole pending result of its storage (an
// NArray left pending by an operator), so its map can be spliced in.
bool whole_pending(const Operand &o) {
  return o.storage->pending_map && std::memcmp(&o.desc, &o.storage->pending_map->out, sizeof(ph_view_desc)) == 0;
}

// Bytes per result element that a pending map's inputs read; an input whose
// buffer holds fewer elements than the result (a broadcast) counts in part.
double pending_read_bytes(const PendingMap &pm) {
  int64_t n = 1;
  for (int k = 0; k < pm.out.view.rank; ++k) n *= pm.out.view.shape[k];
  double b = 0.0;
  for (const Operand &i : pm.inputs)
    b += (double)dtype_size((DType)i.desc.dtype) * std::min(1.0, (double)i.desc.extent / (double)std::max<int64_t>(n, 1));
  return b;
}

// Whether a pending result is spliced into one more consumer.  One whose
// inputs read at most an eighth more bytes per element than the result
// holds always is (splicing costs next to no traffic; a broadcast operand
// adds its small share).  One that reads more (a Bool mask computed from
// Int32 values) is spliced into its first consumer only: a second consumer
// runs it into its buffer and reads that, so a result used many times is
// computed once.
bool splice_pending(PendingMap &pm) {
  if (pm.spliced > 0 && pending_read_bytes(pm) > 1.125 * (double)dtype_size((DType)pm.out.dtype)) return false;
  ++pm.spliced;
  return true;
}

// `fuse`: pending operator results are spliced into the program instead of
// being run first (splice_pending), except those that read `dst` (the
// buffer the map writes): the reference computed them before the write, so
// they stay a snapshot.
MapPlan plan_map(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn,
                 bool fuse, const Storage *dst) {
  if (ops.empty()) throw ArgumentError("map_with needs an operand");
  for (size_t i = 1; i < ops.size(); ++i)
    if (ops[i].shape != ops[0].shape)
      throw ShapeError("Could not simultaneously map MultiIndexables with shapes " + shape_to_s(ops[i].shape) +
                       " and " + shape_to_s(ops[0].shape) + ".");
  MapPlan p;
  std::function<Expr(const Operand &)> leaf = [&](const Operand &o) -> Expr {
    if (fuse && whole_pending(o)) {
      PendingMap &pm = *o.storage->pending_map;
      bool reads_dst = false;
      for (const Operand &i : pm.inputs) reads_dst = reads_dst || i.storage.get() == dst;
      if (!reads_dst && splice_pending(pm)) {
        std::vector<Expr> subs;
        for (const Operand &i : pm.inputs) subs.push_back(leaf(i));
        Expr e = pm.body.substitute(subs);
        return o.body ? o.body->substitute({e}) : e;
      }
    }
    int idx = -1;
    if (o.deep && dst && o.deep->storage.get() == dst) {
      // a view past one descriptor of the buffer written: read through its
      // index array, element by element in the loop's order
      for (size_t k = 0; k < p.inputs.size(); ++k)
        if (p.chain[k] == o.deep.get()) idx = (int)k;
      if (idx < 0) {
        if (p.inputs.size() >= PH_MAX_INPUTS) throw ArgumentError("more than 8 distinct operands in one map");
        Operand in;
        in.storage = o.deep->storage;
        in.desc = o.deep->desc;
        in.desc.view.rank = (int32_t)o.shape.size();
        for (size_t i = 0; i < o.shape.size(); ++i) in.desc.view.shape[i] = o.shape[i];
        in.shape = o.shape;
        p.inputs.push_back(in);
        p.index.push_back(index_of_chain(*o.deep));
        p.chain.push_back(o.deep.get());
        idx = (int)p.inputs.size() - 1;
      }
      Expr e = Expr::input(idx, o.deep->storage->dtype);
      return o.body ? o.body->substitute({e}) : e;
    }
    for (size_t k = 0; k < p.inputs.size(); ++k)
      if (!p.chain[k] && p.inputs[k].storage == o.storage &&
          std::memcmp(&p.inputs[k].desc, &o.desc, sizeof(ph_view_desc)) == 0)
        idx = (int)k;
    if (idx < 0) {
      if (p.inputs.size() >= PH_MAX_INPUTS) throw ArgumentError("more than 8 distinct operands in one map");
      Operand in = o;
      in.body.reset();
      in.deep.reset();  // the gathered copy is what a map not writing its source reads
      int staged = 0;
      for (size_t k = 0; k < p.inputs.size(); ++k) staged += !p.chain[k] && p.inputs[k].desc.nstage > 0 ? 1 : 0;
      if (in.desc.nstage > 0 && staged >= 2) {
        // ph_map takes at most two reshaped/tiled operands: gather this one
        NArray g = materialize_operand(in);
        in = g.operand();
      }
      p.inputs.push_back(in);
      p.index.emplace_back();
      p.chain.push_back(nullptr);
      idx = (int)p.inputs.size() - 1;
    }
    Expr e = Expr::input(idx, (DType)o.desc.dtype);
    return o.body ? o.body->substitute({e}) : e;
  };
  std::vector<Expr> leaves;
  for (const Operand &o : ops) leaves.push_back(leaf(o));
  p.body = fn(leaves);
  return p;
}

// plan_map fusing pending operands where the fused program stays within a
// map's limits (inputs and instructions); otherwise they run first.
MapPlan plan_map(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn,
                 const Storage *dst) {
  bool any = false;
  for (const Operand &o : ops) any = any || whole_pending(o);
  if (any) {
    try {
      MapPlan p = plan_map(ops, fn, true, dst);
      (void)p.body.compile((int)p.inputs.size());
      return p;
    } catch (const ArgumentError &) {
      // too large to fuse: the pending operands run on their own
    }
  }
  return plan_map(ops, fn, false, dst);
}

// Runs the plan into the view `od` of `dst`.  Inputs that read dst's
// buffer see the writes of earlier elements, as the reference's loops do.
void launch_map(const MapPlan &p, const std::shared_ptr<Storage> &dst, const ph_view_desc &od) {
  ph_prog prog = p.body.compile((int)p.inputs.size());
  std::vector<const void *> ptrs;
  std::vector<ph_view_desc> descs;
  std::vector<const int64_t *> index;
  bool indexed = false;
  for (size_t k = 0; k < p.inputs.size(); ++k) {
    ptrs.push_back(p.inputs[k].storage->dev_read());
    descs.push_back(p.inputs[k].desc);
    index.push_back(p.chain[k] ? (const int64_t *)p.index[k].storage->dev_read() : nullptr);
    indexed = indexed || p.chain[k];
  }
  void *op = dst->dev_write();
  int32_t flags = 0;
  if (indexed)
    check(ph_map_indexed(&prog, (int32_t)p.inputs.size(), ptrs.data(), descs.data(), index.data(), op, &od, nullptr,
                         &flags),
          "map");
  else
    check(ph_map(&prog, (int32_t)p.inputs.size(), ptrs.data(), descs.data(), op, &od, &flags), "map");
  check_flags(flags, "elementwise");
}
}  // namespace

// Fused map over several operands of one shape, into `out` (a contiguous
// NArray of the result type), or into a new NArray when out is null.  A new
// result whose program cannot raise is left pending (Storage::pending_map),
// so that the operator that uses it next fuses it into its own kernel.
static NArray map_core(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn,
                       const NArray *out_in) {
  MapPlan p = plan_map(ops, fn, out_in ? out_in->storage.get() : nullptr);
  NArray out;
  if (out_in) {
    out = *out_in;
    if (compute_dtype(out.dtype()) != p.body.dtype() || out.dtype() == DType::BF16)
      throw ArgumentError(std::string("cannot store ") + dtype_name(p.body.dtype()) + " values into a " +
                          dtype_name(out.dtype()) + " array");
  } else {
    out = NArray::empty(ops[0].shape, p.body.dtype());
  }
  if (out.size() == 0) return out;
  const ph_prog prog = p.body.compile((int)p.inputs.size());
  bool uses_index = false;
  for (int k = 0; k < prog.n_insn; ++k) uses_index = uses_index || prog.insn[k].op == PH_OP_INDEX;
  // left pending only when running it later cannot raise: no checked Int32
  // arithmetic, and no element index past Int32 (ph_map refuses that shape)
  if (!out_in && !prog_may_raise(prog) && !(uses_index && out.size() > INT32_MAX)) {
    auto pm = std::make_shared<PendingMap>();
    pm->inputs = p.inputs;
    pm->body = p.body;
    pm->out = out.desc();
    for (const Operand &i : pm->inputs) {
      auto &rd = i.storage->readers;
      rd.erase(std::remove_if(rd.begin(), rd.end(), [](const std::weak_ptr<Storage> &w) { return w.expired(); }),
               rd.end());
      rd.push_back(out.storage);
    }
    out.storage->pending_map = pm;
    return out;
  }
  launch_map(p, out.storage, out.desc());
  return out;
}

// The same into the view `od` of `dst` (a scatter of a lazily evaluated
// source); false when the map cannot store dst's element type.
static bool map_into(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn,
                     const std::shared_ptr<Storage> &dst, const ph_view_desc &od) {
  MapPlan p = plan_map(ops, fn, dst.get());
  if (dst->dtype == DType::BF16 || compute_dtype(dst->dtype) != p.body.dtype()) return false;
  launch_map(p, dst, od);
  return true;
}

NArray map_with(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn) {
  return map_core(ops, fn, nullptr);
}

static void same_shape(const std::string &op, const Operand &a, const Operand &b) {
  if (a.shape != b.shape) {
    std::string m = "The shape of this MultiIndexable (" + shape_to_s(a.shape) +
                    ") does not match the shape of the one provided (" + shape_to_s(b.shape) + "), so '" + op +
                    "' cannot be applied element-wise.";
    if (ShapeUtil::shape_to_size(a.shape) == 1 || ShapeUtil::shape_to_size(b.shape) == 1)
      m.replace(m.size() - 1, 1, ". Did you mean to call to_scalar on one of the arguments?");
    throw ShapeError(m);
  }
}

NArray elementwise(const std::string &op, const Operand &a, const Operand &b) {
  same_shape(op, a, b);
  return map_with({a, b}, [&](const std::vector<Expr> &x) { return Expr::binary(op, x[0], x[1]); });
}

NArray elementwise(const std::string &op, const Operand &a, const Scalar &b) {
  return map_with({a}, [&](const std::vector<Expr> &x) { return Expr::binary(op, x[0], Expr::constant(b)); });
}

NArray elementwise(const std::string &op, const Scalar &a, const Operand &b) {
  return map_with({b}, [&](const std::vector<Expr> &x) { return Expr::binary(op, Expr::constant(a), x[0]); });
}

NArray elementwise_unary(const std::string &op, const Operand &a) {
  return map_with({a}, [&](const std::vector<Expr> &x) { return Expr::unary(op, x[0]); });
}

NArray map(const Operand &a, const std::function<Expr(const Expr &)> &fn) {
  return map_with({a}, [&](const std::vector<Expr> &x) { return fn(x[0]); });
}

NArray &NArray::map_inplace(const std::function<Expr(const Expr &)> &fn) {
  Operand self = operand();
  Expr body = fn(Expr::input(0, dtype()));
  if (body.dtype() != compute_dtype(dtype()))
    throw ArgumentError(std::string("map! cannot change the element type (") + dtype_name(dtype()) + " -> " +
                        dtype_name(body.dtype()) + ")");
  if (dtype() == DType::BF16) throw ArgumentError("map! on BFloat16 storage is not supported");
  if (size() == 0) return *this;
  void *sp = storage->dev_write();
  ph_prog prog = body.compile(1);
  ph_view_desc d = desc();
  const void *ptrs[1] = {sp};
  int32_t flags = 0;
  check(ph_map(&prog, 1, ptrs, &d, sp, &d, &flags), "map!");
  check_flags(flags, "map!");
  return *this;
}

NArray NArray::map_with_index(const std::function<Expr(const Expr &, const Expr &)> &fn) const {
  return map(operand(), [&](const Expr &x) { return fn(x, Expr::index()); });
}

NArray &NArray::map_with_index_inplace(const std::function<Expr(const Expr &, const Expr &)> &fn) {
  return map_inplace([&](const Expr &x) { return fn(x, Expr::index()); });
}

// Coordinates follow from the lexicographic index by floored division and
// modulo with the axis strides of a contiguous array of `shape`: the same
// coordinate LexIterator.cover (src/n_array.cr:643) reaches after idx steps.
std::vector<Expr> coord_exprs(const Shape &shape, const Expr &idx) {
  std::vector<Expr> c(shape.size(), idx);
  int64_t stride = 1;
  for (size_t k = shape.size(); k-- > 0;) {
    Expr q = stride == 1 ? idx : Expr::binary("//", idx, Expr::constant(Scalar::i32(stride)));
    c[k] = k == 0 ? q : Expr::binary("%", q, Expr::constant(Scalar::i32(shape[k])));
    stride *= std::max<int64_t>(shape[k], 1);
    if (stride > INT32_MAX) stride = INT32_MAX;  // only reached by shapes the Int32 index rejects
  }
  return c;
}

NArray map_with_coord(const Operand &a,
                      const std::function<Expr(const Expr &, const std::vector<Expr> &, const Expr &)> &fn) {
  return map(a, [&](const Expr &x) {
    Expr i = Expr::index();
    return fn(x, coord_exprs(a.shape, i), i);
  });
}

NArray &NArray::map_with_coord_inplace(
    const std::function<Expr(const Expr &, const std::vector<Expr> &, const Expr &)> &fn) {
  Shape sh = shape();
  return map_inplace([&](const Expr &x) {
    Expr i = Expr::index();
    return fn(x, coord_exprs(sh, i), i);
  });
}

NArray eq(const Operand &a, const Operand &b) {
  if (a.shape != b.shape)
    throw DimensionError("Cannot compute the element-wise equality between this MultiIndexable (shape: " +
                         shape_to_s(a.shape) + ") and the one provided (shape: " + shape_to_s(b.shape) + ").");
  return map_with({a, b}, [](const std::vector<Expr> &x) { return Expr::binary("==", x[0], x[1]); });
}

// ===========================================================================
// Reductions
// ===========================================================================
DeviceDoubles::DeviceDoubles(int64_t n_) : n(n_) {
  init();
  ptr = (double *)pool_alloc(round_bytes((size_t)std::max<int64_t>(n, 1) * sizeof(double)));
}
DeviceDoubles::~DeviceDoubles() {
  if (ptr) pool_release(ptr, round_bytes((size_t)std::max<int64_t>(n, 1) * sizeof(double)));
}
std::vector<double> DeviceDoubles::to_host() const {
  std::vector<double> h((size_t)n);
  if (n) check(ph_d2h(h.data(), ptr, (size_t)n * sizeof(double)), "result copy");
  return h;
}

struct Prepared {
  ph_prog prog;
  std::vector<const void *> ptrs;
  std::vector<ph_view_desc> descs;
  DType vdt;
};

static Prepared prepare(const Operand &a) {
  Prepared p;
  Expr e = Expr::input(0, (DType)a.desc.dtype);
  if (a.body) e = a.body->substitute({e});
  if (e.dtype() == DType::Bool) throw ArgumentError("undefined method '+' for Bool (sum of a Bool MultiIndexable)");
  p.vdt = e.dtype();
  p.prog = e.compile(1);
  p.ptrs.push_back(a.storage->dev_read());
  p.descs.push_back(a.desc);
  return p;
}

static int64_t count_of(const Operand &a) { return ShapeUtil::shape_to_size(a.shape); }

Scalar sum(const Operand &a, Acc acc) {
  if (count_of(a) == 0) {
    Expr e = Expr::input(0, (DType)a.desc.dtype);
    if (a.body) e = a.body->substitute({e});
    return e.dtype() == DType::I32 ? Scalar::i32(0) : Scalar::f64(0.0);
  }
  Prepared p = prepare(a);
  DeviceDoubles out(1);
  int32_t flags = 0;
  check(ph_reduce(&p.prog, 1, p.ptrs.data(), p.descs.data(), -1, PH_RED_SUM, (int32_t)acc, out.ptr, nullptr, &flags),
        "sum");
  check_flags(flags, "sum");
  double v = out.to_host()[0];
  if (p.vdt == DType::I32) return Scalar::i32((int64_t)v);
  return Scalar::f64(v);
}

Scalar max(const Operand &a) {
  if (count_of(a) == 0) throw EmptyError("Empty enumerable");
  Prepared p = prepare(a);
  DeviceDoubles out(1);
  int32_t flags = 0;
  check(ph_reduce(&p.prog, 1, p.ptrs.data(), p.descs.data(), -1, PH_RED_MAX, PH_ACC_F64, nullptr, out.ptr, &flags),
        "max");
  double v = out.to_host()[0];
  if (p.vdt == DType::I32) return Scalar::i32((int64_t)v);
  if (p.vdt == DType::F32) return Scalar::f32(v);
  return Scalar::f64(v);
}

std::vector<double> slices_sum(const Operand &a, int axis, Acc acc) {
  if (axis < 0 || axis >= (int)a.shape.size()) throw IndexError("axis out of range");
  const int64_t n = a.shape[(size_t)axis];
  if (n == 0) return {};
  if (count_of(a) == 0) return std::vector<double>((size_t)n, 0.0);
  Prepared p = prepare(a);
  DeviceDoubles out(n);
  int32_t flags = 0;
  check(ph_reduce(&p.prog, 1, p.ptrs.data(), p.descs.data(), axis, PH_RED_SUM, (int32_t)acc, out.ptr, nullptr, &flags),
        "slices sum");
  check_flags(flags, "slices sum");
  return out.to_host();
}

std::vector<double> slices_max(const Operand &a, int axis) {
  if (axis < 0 || axis >= (int)a.shape.size()) throw IndexError("axis out of range");
  const int64_t n = a.shape[(size_t)axis];
  if (n == 0) return {};
  if (count_of(a) == 0) throw EmptyError("Empty enumerable");
  Prepared p = prepare(a);
  DeviceDoubles out(n);
  int32_t flags = 0;
  check(ph_reduce(&p.prog, 1, p.ptrs.data(), p.descs.data(), axis, PH_RED_MAX, PH_ACC_F64, nullptr, out.ptr, &flags),
        "slices max");
  return out.to_host();
}

Reduce2D reduce2d_all(const Operand &a, Acc acc) {
  if (a.shape.size() != 2) throw DimensionError("reduce2d_all needs a 2-dimensional MultiIndexable");
  if (count_of(a) == 0) throw EmptyError("Empty enumerable");
  Prepared p = prepare(a);
  const int64_t R = a.shape[0], C = a.shape[1];
  DeviceDoubles buf(2 + 2 * R + 2 * C);
  int32_t flags = 0;
  double *b = buf.ptr;
  check(ph_reduce2d_all(&p.prog, 1, p.ptrs.data(), p.descs.data(), (int32_t)acc, b, b + 1, b + 2, b + 2 + R,
                        b + 2 + 2 * R, b + 2 + 2 * R + C, &flags),
        "reduce2d_all");
  check_flags(flags, "reduce2d_all");
  std::vector<double> h = buf.to_host();
  Reduce2D r;
  r.full_sum = h[0];
  r.full_max = h[1];
  r.row_sum.assign(h.begin() + 2, h.begin() + 2 + R);
  r.row_max.assign(h.begin() + 2 + R, h.begin() + 2 + 2 * R);
  r.col_sum.assign(h.begin() + 2 + 2 * R, h.begin() + 2 + 2 * R + C);
  r.col_max.assign(h.begin() + 2 + 2 * R + C, h.end());
  return r;
}

NArray reduce2d_all_device(const Operand &a, Acc acc) {
  if (a.shape.size() != 2) throw DimensionError("reduce2d_all needs a 2-dimensional MultiIndexable");
  if (count_of(a) == 0) throw EmptyError("Empty enumerable");
  Prepared p = prepare(a);
  const int64_t R = a.shape[0], C = a.shape[1];
  NArray out = NArray::empty({2 + 2 * R + 2 * C}, DType::F64);
  double *b = (double *)out.storage->dev_write();
  int32_t flags = 0;
  check(ph_reduce2d_all(&p.prog, 1, p.ptrs.data(), p.descs.data(), (int32_t)acc, b, b + 1, b + 2, b + 2 + R,
                        b + 2 + 2 * R, b + 2 + 2 * R + C, &flags),
        "reduce2d_all");
  check_flags(flags, "reduce2d_all");
  return out;
}

bool equals(const Operand &a, const Operand &b) {
  if (a.shape != b.shape) return false;
  if (count_of(a) == 0) return true;
  NArray ne = map_with({a, b}, [](const std::vector<Expr> &x) {
    return Expr::unary("to_i32", Expr::binary("!=", x[0], x[1]));
  });
  Scalar s = sum(ne.operand());
  return s.i == 0;
}

// ===========================================================================
// Heat
// ===========================================================================
std::vector<double> heat2d(NArray &u, double r, int64_t steps, int64_t every) {
  if (u.dimensions() != 2) throw DimensionError("heat2d needs a 2-dimensional NArray");
  if (u.dtype() != DType::F32 && u.dtype() != DType::F64) throw ArgumentError("heat2d needs Float32 or Float64");
  NArray tmp = NArray::empty(u.shape(), u.dtype());
  const int64_t ne = every > 0 ? steps / every + 1 : 0;
  DeviceDoubles e(std::max<int64_t>(ne, 1));
  check(ph_heat2d((int32_t)u.dtype(), u.device_ptr(), tmp.device_ptr(), u.shape()[0], u.shape()[1], r, steps,
                  every, every > 0 ? e.ptr : nullptr),
        "heat2d");
  check(ph_sync(), "heat2d sync");
  if (ne == 0) return {};
  std::vector<double> h = e.to_host();
  h.resize((size_t)ne);
  return h;
}

// ===========================================================================
// Batched small arrays
// ===========================================================================
Batch::Batch(const std::vector<Shape> &shapes_, const std::vector<float> &scalars_)
    : shapes(shapes_), scalars(scalars_) {
  init();
  if (shapes.size() != scalars.size()) throw ArgumentError("one scalar per array");
  if (shapes.size() > (size_t)INT32_MAX) throw ArgumentError("too many arrays");
  const int32_t n = (int32_t)shapes.size();
  std::vector<int32_t> ranks((size_t)n);
  std::vector<int64_t> flat((size_t)n * 3, 0);
  for (int32_t k = 0; k < n; ++k) {
    const Shape &s = shapes[(size_t)k];
    if (s.empty() || s.size() > 3) throw DimensionError("batched arrays have rank 1, 2 or 3");
    for (int64_t d : s)
      if (d < 0) throw DimensionError("Cannot create NArray: One or more of the provided dimensions was negative.");
    ranks[(size_t)k] = (int32_t)s.size();
    for (size_t i = 0; i < s.size(); ++i) flat[(size_t)k * 3 + i] = s[i];
  }
  check(ph_batch_create(&handle_, n, ranks.data(), flat.data(), nullptr, scalars.data(), nullptr, nullptr, nullptr),
        "batch");
  int64_t nxy = 0, nz = 0;
  offsets_x.resize((size_t)n);
  offsets_z.resize((size_t)n);
  check(ph_batch_sizes(handle_, &nxy, &nz, offsets_x.data(), offsets_z.data()), "batch sizes");
  x = NArray::empty({nxy}, DType::F32);
  y = NArray::empty({nxy}, DType::F32);
  z = NArray::empty({nz}, DType::F32);
  sums_dev = std::make_unique<DeviceDoubles>(std::max<int64_t>(n, 1));
  check(ph_batch_bind(handle_, (const float *)x.device_ptr(), (float *)y.device_ptr(), (float *)z.device_ptr(),
                      sums_dev->ptr),
        "batch bind");
}

Batch::~Batch() {
  if (handle_) ph_batch_destroy(handle_);
}

NArray Batch::piece(const NArray &flat, int64_t off, const Shape &shape) const {
  NArray r = NArray::empty(shape, DType::F32);
  const int64_t n = r.size();
  if (n)
    check(ph_d2d(r.device_ptr(), (const float *)flat.storage->dev_read() + off, (size_t)n * 4), "batch element copy");
  return r;
}

static void check_k(int64_t k, int64_t n) {
  if (k < 0 || k >= n) throw IndexError("batch index out of range");
}

NArray Batch::x_at(int64_t k) const {
  check_k(k, count());
  return piece(x, offsets_x[(size_t)k], shapes[(size_t)k]);
}
NArray Batch::y_at(int64_t k) const {
  check_k(k, count());
  return piece(y, offsets_x[(size_t)k], shapes[(size_t)k]);
}
NArray Batch::z_at(int64_t k) const {
  check_k(k, count());
  Shape zs = shapes[(size_t)k];
  for (int64_t &d : zs) d = (d + 1) / 2;
  return piece(z, offsets_z[(size_t)k], zs);
}

void Batch::fill_uniform(uint64_t seed, double a, double b) {
  check(ph_fill_random(x.device_ptr(), PH_F32, x.size(), seed, 0, 0, a, b), "batch random fill");
}

// The kernels use the buffers bound at construction: bring x's device copy up
// to date and drop y's and z's host mirrors before each run.
void Batch::prepare_run() {
  x.storage->dev_read();
  y.storage->dev_write();
  z.storage->dev_write();
}

void Batch::run(int mode) {
  prepare_run();
  check(ph_batch_run(handle_, mode), "batch run");
}

std::tuple<int64_t, int, int> Batch::tune() {
  int64_t chunk = 0;
  int32_t grid = 0, dynamic = 0;
  prepare_run();
  check(ph_batch_tune(handle_, &chunk, &grid, &dynamic), "batch tune");
  return {chunk, (int)grid, (int)dynamic};
}

void Batch::schedule(int64_t chunk, int grid, int dynamic) {
  check(ph_batch_schedule(handle_, chunk, grid, dynamic), "batch schedule");
}

std::vector<double> Batch::sums() const {
  std::vector<double> h = sums_dev->to_host();
  h.resize((size_t)count());
  return h;
}

GBatch::GBatch(DType x_dtype_, DType y_dtype_, const std::vector<Expr> &progs, const std::vector<GBatchRecord> &recs)
    : x_dtype(x_dtype_), y_dtype(y_dtype_), records(recs) {
  init();
  if (records.size() > (size_t)INT32_MAX) throw ArgumentError("too many arrays");
  std::vector<ph_prog> ps;
  for (const Expr &e : progs) ps.push_back(e.compile(2));
  std::vector<ph_batch_rec> rs(records.size());
  for (size_t k = 0; k < records.size(); ++k) {
    const GBatchRecord &g = records[k];
    ph_batch_rec &r = rs[k];
    std::memset(&r, 0, sizeof(r));
    if (g.shape.empty() || g.shape.size() > 3) throw DimensionError("batched arrays have rank 1, 2 or 3");
    for (int64_t d : g.shape)
      if (d < 0) throw DimensionError("Cannot create NArray: One or more of the provided dimensions was negative.");
    r.rank = (int32_t)g.shape.size();
    for (size_t i = 0; i < g.shape.size(); ++i) r.shape[i] = g.shape[i];
    r.map = g.map;
    r.red = g.red;
    r.red_ops = g.red >= 0 ? g.red_ops : 0;
    r.scalar = g.scalar;
    Shape zs;
    if (g.slice) {
      // the region as `view` canonicalises it (IndexRegion.new(literal, shape))
      const IndexRegion ir = IndexRegion::from_literal(g.region, g.shape);
      r.slice = 1;
      for (size_t i = 0; i < g.shape.size(); ++i) {
        r.first[i] = ir.first[i];
        r.step[i] = ir.step[i];
        r.count[i] = ir.proper_shape[i];
      }
      zs = ir.reduced_shape;
    }
    z_shapes.push_back(zs);
  }
  check(ph_gbatch_create(&handle_, (int32_t)x_dtype, (int32_t)y_dtype, (int32_t)ps.size(), ps.data(),
                         (int32_t)rs.size(), rs.data()),
        "batch");
  int64_t nx = 0, nz = 0;
  offsets_xy.resize(records.size());
  offsets_z.resize(records.size());
  check(ph_gbatch_sizes(handle_, &nx, &nz, offsets_xy.data(), offsets_z.data()), "batch sizes");
  x = NArray::empty({nx}, x_dtype);
  y = NArray::empty({nx}, y_dtype);
  z = NArray::empty({nz}, x_dtype);
  sums_dev = std::make_unique<DeviceDoubles>(std::max<int64_t>(count(), 1));
  maxes_dev = std::make_unique<DeviceDoubles>(std::max<int64_t>(count(), 1));
}

GBatch::~GBatch() {
  if (handle_) ph_gbatch_destroy(handle_);
}

NArray GBatch::piece(const NArray &flat, DType dt, int64_t off, const Shape &shape) const {
  NArray r = NArray::empty(shape, dt);
  const int64_t n = r.size();
  const size_t es = dtype_size(dt);
  if (n)
    check(ph_d2d(r.device_ptr(), (const char *)flat.storage->dev_read() + (size_t)off * es, (size_t)n * es),
          "batch element copy");
  return r;
}

NArray GBatch::x_at(int64_t k) const {
  check_k(k, count());
  return piece(x, x_dtype, offsets_xy[(size_t)k], records[(size_t)k].shape);
}
NArray GBatch::y_at(int64_t k) const {
  check_k(k, count());
  return piece(y, y_dtype, offsets_xy[(size_t)k], records[(size_t)k].shape);
}
NArray GBatch::z_at(int64_t k) const {
  check_k(k, count());
  if (!records[(size_t)k].slice) throw ArgumentError("batch record has no region");
  return piece(z, x_dtype, offsets_z[(size_t)k], z_shapes[(size_t)k]);
}

void GBatch::set_x(const NArray &flat) {
  if (flat.dtype() != x_dtype || flat.size() != x.size()) throw ArgumentError("set_x: wrong dtype or size");
  if (x.size())
    check(ph_d2d(x.storage->dev_write(), flat.storage->dev_read(), (size_t)x.size() * dtype_size(x_dtype)),
          "batch set_x");
}

void GBatch::prepare_run() {
  x.storage->dev_read();
  y.storage->dev_write();
  z.storage->dev_write();
}

void GBatch::run(int mode) {
  prepare_run();
  int32_t flags = 0;
  check(ph_gbatch_run(handle_, x.storage->dev_read(), y.storage->dev_write(), z.storage->dev_write(), sums_dev->ptr,
                      maxes_dev->ptr, mode, &flags),
        "batch run");
  check_flags(flags, "batch");
}

std::tuple<int64_t, int, int> GBatch::tune() {
  int64_t chunk = 0;
  int32_t grid = 0, dynamic = 0;
  prepare_run();
  check(ph_gbatch_tune(handle_, x.storage->dev_read(), y.storage->dev_write(), z.storage->dev_write(), sums_dev->ptr,
                       maxes_dev->ptr, &chunk, &grid, &dynamic),
        "batch tune");
  return {chunk, (int)grid, (int)dynamic};
}

std::vector<double> GBatch::sums() const {
  std::vector<double> h = sums_dev->to_host();
  h.resize((size_t)count());
  return h;
}

std::vector<double> GBatch::maxes() const {
  std::vector<double> h = maxes_dev->to_host();
  h.resize((size_t)count());
  return h;
}

}  // namespace phase
