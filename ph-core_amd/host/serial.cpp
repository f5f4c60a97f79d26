// NArray serialisation: #to_json / .from_json / #to_yaml / .from_yaml
// (src/n_array.cr:807-911).  The document is {"shape": [...], "elements":
// [...]} with the elements in buffer (lexicographic) order; element text is
// the element type's to_s.  The device buffer is copied to or from the host
// once per call.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include "phase.hpp"

namespace phase {

// ---------------------------------------------------------------------------
// Float text: the shortest decimal digits that read back to the same value,
// laid out as Crystal prints floats: plain notation while the decimal point
// falls within 16 digits left and 3 zeros right of the first digit, else
// d.ddde+XX / d.ddde-XX; always with a fractional part ("1.0").
// ---------------------------------------------------------------------------
void append_float(std::string &out, double v, bool single) {
  if (std::isnan(v)) {
    out += "NaN";
    return;
  }
  if (std::signbit(v)) {
    out += '-';
    v = -v;
  }
  if (std::isinf(v)) {
    out += "Infinity";
    return;
  }
  if (v == 0.0) {
    out += "0.0";
    return;
  }
  char buf[64];
  const std::to_chars_result r = single
                                     ? std::to_chars(buf, buf + sizeof(buf), (float)v, std::chars_format::scientific)
                                     : std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
  // buf: d[.ddd]e(+|-)XX
  char digits[40];
  int length = 0;
  const char *p = buf;
  for (; p < r.ptr && *p != 'e'; ++p)
    if (*p != '.') digits[length++] = *p;
  int e10 = 0;
  std::from_chars(p + 2, r.ptr, e10);
  if (p[1] == '-') e10 = -e10;
  int point = e10 + 1;  // position of the decimal point after the first digit
  const int exp = point;
  const bool exp_mode = point > 15 || point < -3;
  if (exp_mode) point = 1;
  if (point < 1) out += '0';
  int i = 0;
  if (point > 0) {
    for (; i < length && i < point; ++i) out += digits[i];
    for (int k = i; k < point; ++k) out += '0';
  }
  out += '.';
  for (int k = 0; k < -point; ++k) out += '0';
  out.append(digits + i, (size_t)(length - i));
  if (point >= length) out += '0';
  if (exp_mode) {
    out += 'e';
    out += exp > 0 ? '+' : '-';
    char eb[16];
    const std::to_chars_result er = std::to_chars(eb, eb + sizeof(eb), std::abs(exp - 1));
    out.append(eb, (size_t)(er.ptr - eb));
  }
}

std::string float_to_s(double v, bool single) {
  std::string out;
  append_float(out, v, single);
  return out;
}

namespace {

// Threads that format or parse n elements: one per 16384 elements, at most
// 16 and at most the host's hardware threads.
int chunk_threads(int64_t n) {
  int64_t t = std::min<int64_t>(16, std::max<int64_t>(1, n / 16384));
  const unsigned hc = std::thread::hardware_concurrency();
  if (hc) t = std::min<int64_t>(t, (int64_t)hc);
  return (int)t;
}

// f(t, begin, end) over T consecutive chunks of [0, n), chunk 0 on the
// calling thread.  f must not throw: it records failures for the caller.
template <typename F>
void for_chunks(int64_t n, int T, F f) {
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back([=, &f] { f(t, n * t / T, n * (t + 1) / T); });
  f(0, 0, n / T);
  for (std::thread &th : pool) th.join();
}

// Element i of the host copy h (element type dt) as its type's to_s
// (YAML spells the non-finite floats .inf, -.inf, .nan).
void append_element(std::string &out, const unsigned char *h, int64_t i, DType dt, bool yaml) {
  auto flt = [&](double v, bool single) {
    if (yaml && std::isinf(v)) out += v > 0 ? ".inf" : "-.inf";
    else if (yaml && std::isnan(v)) out += ".nan";
    else append_float(out, v, single);
  };
  switch (dt) {
    case DType::I32: {
      int32_t v;
      std::memcpy(&v, h + (size_t)i * 4, 4);
      char b[16];
      const std::to_chars_result r = std::to_chars(b, b + sizeof(b), v);
      out.append(b, (size_t)(r.ptr - b));
      break;
    }
    case DType::F32: {
      float v;
      std::memcpy(&v, h + (size_t)i * 4, 4);
      flt(v, true);
      break;
    }
    case DType::F64: {
      double v;
      std::memcpy(&v, h + (size_t)i * 8, 8);
      flt(v, false);
      break;
    }
    case DType::BF16: {
      uint16_t b;
      std::memcpy(&b, h + (size_t)i * 2, 2);
      const uint32_t u = (uint32_t)b << 16;
      float v;
      std::memcpy(&v, &u, 4);
      flt(v, true);
      break;
    }
    case DType::Bool:
      out += h[(size_t)i] ? "true" : "false";
      break;
  }
}

// Whether element i is a float that JSON cannot hold (NaN, +-Infinity).
bool non_finite(const unsigned char *h, int64_t i, DType dt) {
  switch (dt) {
    case DType::F32: {
      float v;
      std::memcpy(&v, h + (size_t)i * 4, 4);
      return !std::isfinite(v);
    }
    case DType::F64: {
      double v;
      std::memcpy(&v, h + (size_t)i * 8, 8);
      return !std::isfinite(v);
    }
    case DType::BF16: {
      uint16_t b;
      std::memcpy(&b, h + (size_t)i * 2, 2);
      return (b & 0x7f80) == 0x7f80;
    }
    default:
      return false;
  }
}

std::vector<unsigned char> host_copy(const NArray &a) {
  std::vector<unsigned char> h((size_t)a.size() * dtype_size(a.dtype()));
  if (!h.empty()) a.to_host(h.data(), h.size());
  return h;
}

// Elements of `a` (any dtype) as text, in buffer order.
std::vector<std::string> element_texts(const NArray &a, bool yaml) {
  const std::vector<unsigned char> h = host_copy(a);
  std::vector<std::string> t((size_t)a.size());
  for (int64_t i = 0; i < a.size(); ++i) append_element(t[(size_t)i], h.data(), i, a.dtype(), yaml);
  return t;
}

// Parses one element text into slot i of h; false when it is not a value of
// `dtype`.
bool parse_element(std::string_view s, DType dtype, unsigned char *h, size_t i) {
  switch (dtype) {
    case DType::I32: {
      int64_t v = 0;
      auto r = std::from_chars(s.data(), s.data() + s.size(), v);
      if (r.ec != std::errc() || r.ptr != s.data() + s.size() || v < INT32_MIN || v > INT32_MAX) return false;
      const int32_t w = (int32_t)v;
      std::memcpy(&h[i * 4], &w, 4);
      return true;
    }
    case DType::F32:
    case DType::F64:
    case DType::BF16: {
      double v = 0;
      if (s == ".inf" || s == "+.inf" || s == "Infinity") v = INFINITY;
      else if (s == "-.inf" || s == "-Infinity") v = -INFINITY;
      else if (s == ".nan" || s == "NaN") v = NAN;
      else {
        auto r = std::from_chars(s.data(), s.data() + s.size(), v);
        if (r.ec != std::errc() || r.ptr != s.data() + s.size()) return false;
      }
      if (dtype == DType::F64) {
        std::memcpy(&h[i * 8], &v, 8);
      } else {
        const float f = (float)v;
        if (dtype == DType::F32) {
          std::memcpy(&h[i * 4], &f, 4);
        } else {  // BF16 storage: round to nearest even
          uint32_t u;
          std::memcpy(&u, &f, 4);
          const uint16_t b = std::isnan(f) ? (uint16_t)((u >> 16) | 0x40) : (uint16_t)((u + 0x7fff + ((u >> 16) & 1)) >> 16);
          std::memcpy(&h[i * 2], &b, 2);
        }
      }
      return true;
    }
    case DType::Bool:
      if (s == "true") h[i] = 1;
      else if (s == "false") h[i] = 0;
      else return false;
      return true;
  }
  return false;
}

// Builds an NArray of `dtype` from n element texts, text(i) giving the i-th;
// an invalid text raises for the first one in order.
template <typename Text>
NArray build(const Shape &shape, DType dtype, size_t n, Text text, const char *fmt) {
  std::vector<unsigned char> h(n * dtype_size(dtype));
  const int T = chunk_threads((int64_t)n);
  std::vector<int64_t> bad((size_t)T, -1);
  for_chunks((int64_t)n, T, [&](int t, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      if (!parse_element(text((size_t)i), dtype, h.data(), (size_t)i)) {
        bad[(size_t)t] = i;
        return;
      }
  });
  for (int64_t k : bad)
    if (k >= 0)
      throw ParseError(std::string("Could not read NArray from ") + fmt + ": invalid " + dtype_name(dtype) +
                       " element '" + std::string(text((size_t)k)) + "'");
  return NArray::of_buffer(shape, dtype, h.data(), h.size());
}

void check_count(const Shape &shape, size_t n, const char *fmt) {
  if ((size_t)ShapeUtil::shape_to_size(shape) != n)
    throw ParseError(std::string("Could not read NArray from ") + fmt + ": wrong number of elements for shape " +
                     shape_to_s(shape));
}

// --- a small JSON reader: the subset an NArray document uses, other values skipped
struct Json {
  const std::string &s;
  size_t i = 0;
  explicit Json(const std::string &t) : s(t) {}
  [[noreturn]] void fail(const std::string &what) {
    throw ParseError("Could not read NArray from JSON: " + what + " at character " + std::to_string(i));
  }
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail(std::string("expected '") + c + "'");
  }
  std::string string() {
    ws();
    if (i >= s.size() || s[i] != '"') fail("expected a string");
    ++i;
    std::string r;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\' && i + 1 < s.size()) ++i;
      r += s[i++];
    }
    if (i >= s.size()) fail("unterminated string");
    ++i;
    return r;
  }
  std::string atom() {  // number, true, false, null
    ws();
    const size_t b = i;
    while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '-' || s[i] == '+' || s[i] == '.')) ++i;
    if (i == b) fail("expected a value");
    return s.substr(b, i - b);
  }
  std::vector<std::string> array() {
    std::vector<std::string> r;
    for (const auto &sp : array_spans()) r.push_back(s.substr(sp.first, sp.second));
    return r;
  }
  // an array of numbers or booleans, as (offset, length) spans of the text
  std::vector<std::pair<size_t, size_t>> array_spans() {
    std::vector<std::pair<size_t, size_t>> r;
    expect('[');
    if (eat(']')) return r;
    do {
      ws();
      if (i < s.size() && (s[i] == '[' || s[i] == '{' || s[i] == '"')) fail("expected a number or boolean");
      const size_t b = i;
      while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '-' || s[i] == '+' || s[i] == '.')) ++i;
      if (i == b) fail("expected a value");
      r.emplace_back(b, i - b);
    } while (eat(','));
    expect(']');
    return r;
  }
  void skip() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    if (s[i] == '"') {
      string();
    } else if (s[i] == '[' || s[i] == '{') {
      const char open = s[i], close = open == '[' ? ']' : '}';
      ++i;
      if (eat(close)) return;
      do {
        if (open == '{') {
          string();
          expect(':');
        }
        skip();
      } while (eat(','));
      expect(close);
    } else {
      atom();
    }
  }
};

Shape to_shape(const std::vector<std::string> &t, const char *fmt) {
  Shape s;
  for (const std::string &x : t) {
    int64_t v = 0;
    auto r = std::from_chars(x.data(), x.data() + x.size(), v);
    if (r.ec != std::errc() || r.ptr != x.data() + x.size() || v < INT32_MIN || v > INT32_MAX)
      throw ParseError(std::string("Could not read NArray from ") + fmt + ": invalid shape entry '" + x + "'");
    s.push_back(v);
  }
  return s;
}

}  // namespace

// NArray#to_json (n_array.cr:807-818)
// (the elements are formatted in consecutive chunks on several threads and
// joined in order; a non-finite float raises for the first one in order)
std::string to_json(const NArray &a) {
  std::string out = "{\"shape\":[";
  for (size_t i = 0; i < a.shape().size(); ++i) out += (i ? "," : "") + std::to_string(a.shape()[i]);
  out += "],\"elements\":[";
  const std::vector<unsigned char> h = host_copy(a);
  const int64_t n = a.size();
  const DType dt = a.dtype();
  const int T = chunk_threads(n);
  std::vector<std::string> part((size_t)T);
  std::vector<int64_t> bad((size_t)T, -1);
  for_chunks(n, T, [&](int t, int64_t b, int64_t e) {
    std::string &o = part[(size_t)t];
    o.reserve((size_t)(e - b) * (dtype_size(dt) >= 4 ? 22 : 8));
    for (int64_t i = b; i < e; ++i) {
      if (non_finite(h.data(), i, dt)) {
        bad[(size_t)t] = i;
        return;
      }
      if (i) o += ',';
      append_element(o, h.data(), i, dt, false);
    }
  });
  for (int64_t k : bad)
    if (k >= 0) {
      std::string e;
      append_element(e, h.data(), k, dt, false);
      throw ParseError(e + " not allowed in JSON");
    }
  size_t len = out.size() + 2;
  for (const std::string &p : part) len += p.size();
  out.reserve(len);
  for (const std::string &p : part) out += p;
  return out + "]}";
}

// NArray(T).from_json / NArray.new(pull : JSON::PullParser) (n_array.cr:820-850)
NArray narray_from_json(const std::string &text, DType dtype) {
  Json j(text);
  bool found_shape = false, found_elements = false;
  std::vector<std::string> shape_t;
  std::vector<std::pair<size_t, size_t>> el;
  j.expect('{');
  if (!j.eat('}')) {
    do {
      const std::string key = j.string();
      j.expect(':');
      if (key == "shape") {
        found_shape = true;
        shape_t = j.array();
      } else if (key == "elements") {
        found_elements = true;
        el = j.array_spans();
      } else {
        j.skip();
      }
    } while (j.eat(','));
    j.expect('}');
  }
  j.ws();
  if (j.i != text.size()) j.fail("unexpected trailing characters");
  if (!(found_shape && found_elements))
    throw ParseError("Could not read NArray from JSON: 'shape' and/or 'elements' were missing.");
  const Shape shape = to_shape(shape_t, "JSON");
  check_count(shape, el.size(), "JSON");
  const std::string_view tv(text);
  return build(shape, dtype, el.size(), [&](size_t k) { return tv.substr(el[k].first, el[k].second); }, "JSON");
}

// NArray#to_yaml (n_array.cr:852-866), in the layout of a YAML emitter with
// an 80-column line width: flow sequences for numbers (continued on a new
// line indented by 2 once a line has passed column 80), a block sequence for
// Bool elements.
std::string to_yaml(const NArray &a) {
  std::string out = "---\nshape: [";
  for (size_t i = 0; i < a.shape().size(); ++i) out += (i ? ", " : "") + std::to_string(a.shape()[i]);
  out += "]\n";
  const std::vector<std::string> el = element_texts(a, true);
  if (a.dtype() == DType::Bool && !el.empty()) {
    out += "elements:\n";
    for (const std::string &e : el) out += "- " + e + "\n";
    return out;
  }
  std::string line = "elements: [";
  for (size_t i = 0; i < el.size(); ++i) {
    if (i) {
      line += ',';
      if (line.size() > 80) {
        out += line + "\n";
        line = "  ";
      } else {
        line += ' ';
      }
    }
    line += el[i];
  }
  out += line + "]\n";
  return out;
}

// NArray(T).from_yaml / NArray.new(ctx, node) (n_array.cr:868-911), for the
// documents to_yaml writes: a mapping with `shape` and `elements`, each a
// flow sequence (possibly over several lines) or a block sequence.
NArray narray_from_yaml(const std::string &text, DType dtype) {
  std::vector<std::string> lines;
  {
    size_t b = 0;
    while (b <= text.size()) {
      size_t e = text.find('\n', b);
      if (e == std::string::npos) e = text.size();
      lines.push_back(text.substr(b, e - b));
      b = e + 1;
    }
  }
  auto trim = [](std::string s) {
    size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
  };
  auto split_flow = [&](const std::string &body) {  // body without [ ]
    std::vector<std::string> r;
    std::string cur;
    for (char c : body) {
      if (c == ',') {
        r.push_back(trim(cur));
        cur.clear();
      } else {
        cur += c;
      }
    }
    if (!trim(cur).empty() || !r.empty()) r.push_back(trim(cur));
    return r;
  };
  bool found_shape = false, found_elements = false, mapping = false;
  std::vector<std::string> shape_t, el;
  size_t k = 0;
  while (k < lines.size()) {
    std::string l = trim(lines[k]);
    if (l.empty() || l == "---" || l == "..." || l[0] == '#') {
      ++k;
      continue;
    }
    const size_t colon = l.find(':');
    if (colon == std::string::npos || lines[k][0] == ' ' || lines[k][0] == '-')
      throw ParseError("Could not read NArray from YAML: Expected mapping, found " +
                       std::string(l[0] == '-' || l[0] == '[' ? "YAML::Nodes::Sequence" : "YAML::Nodes::Scalar"));
    mapping = true;
    const std::string key = trim(l.substr(0, colon));
    std::string rest = trim(l.substr(colon + 1));
    std::vector<std::string> seq;
    bool is_seq = true;
    ++k;
    if (!rest.empty() && rest[0] == '[') {  // flow sequence, maybe continued
      std::string body = rest.substr(1);
      while (body.find(']') == std::string::npos && k < lines.size()) body += " " + trim(lines[k++]);
      const size_t close = body.find(']');
      if (close == std::string::npos) throw ParseError("Could not read NArray from YAML: unterminated sequence");
      seq = split_flow(body.substr(0, close));
    } else if (rest.empty()) {  // block sequence on the following lines
      while (k < lines.size()) {
        const std::string m = trim(lines[k]);
        if (m.size() >= 1 && m[0] == '-' && (m.size() == 1 || m[1] == ' ')) {
          seq.push_back(trim(m.substr(1)));
          ++k;
        } else if (m.empty()) {
          ++k;
        } else {
          break;
        }
      }
      if (seq.empty()) is_seq = false;
    } else {
      is_seq = false;
    }
    if (key == "shape") {
      if (!is_seq) throw ParseError("Could not read NArray from YAML: invalid shape");
      found_shape = true;
      shape_t = seq;
    } else if (key == "elements") {
      if (!is_seq)
        throw ParseError("Could not read NArray from YAML: Expected sequence, found YAML::Nodes::Scalar");
      found_elements = true;
      el = seq;
    }
  }
  if (!mapping) throw ParseError("Could not read NArray from YAML: Expected mapping, found YAML::Nodes::Scalar");
  if (!(found_shape && found_elements))
    throw ParseError("Could not read NArray from YAML: 'shape' and/or 'elements' were missing.");
  const Shape shape = to_shape(shape_t, "YAML");
  check_count(shape, el.size(), "YAML");
  return build(shape, dtype, el.size(), [&](size_t k) { return std::string_view(el[k]); }, "YAML");
}

}  // namespace phase
