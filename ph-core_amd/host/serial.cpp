// NArray serialisation: #to_json / .from_json / #to_yaml / .from_yaml
// (src/n_array.cr:807-911).  The document is {"shape": [...], "elements":
// [...]} with the elements in buffer (lexicographic) order; element text is
// the element type's to_s.  The device buffer is copied to or from the host
// once per call.
#include <charconv>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "phase.hpp"

namespace phase {

// ---------------------------------------------------------------------------
// Float text: the shortest decimal digits that read back to the same value,
// laid out as Crystal prints floats: plain notation while the decimal point
// falls within 16 digits left and 3 zeros right of the first digit, else
// d.ddde+XX / d.ddde-XX; always with a fractional part ("1.0").
// ---------------------------------------------------------------------------
std::string float_to_s(double v, bool single) {
  if (std::isnan(v)) return "NaN";
  std::string out;
  if (std::signbit(v)) {
    out += '-';
    v = -v;
  }
  if (std::isinf(v)) return out + "Infinity";
  if (v == 0.0) return out + "0.0";
  char buf[64];
  std::to_chars_result r = single ? std::to_chars(buf, buf + sizeof(buf), (float)v, std::chars_format::scientific)
                                  : std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
  *r.ptr = 0;
  // buf: d[.ddd]e(+|-)XX
  std::string digits;
  const char *p = buf;
  for (; *p && *p != 'e'; ++p)
    if (*p != '.') digits += *p;
  const int e10 = std::atoi(p + 1);
  const int length = (int)digits.size();
  int point = e10 + 1;  // position of the decimal point after the first digit
  const int exp = point;
  const bool exp_mode = point > 15 || point < -3;
  if (exp_mode) point = 1;
  if (point < 1) out += '0';
  int i = 0;
  if (point > 0) {
    for (; i < length && i < point; ++i) out += digits[(size_t)i];
    for (int k = i; k < point; ++k) out += '0';
  }
  out += '.';
  for (int k = 0; k < -point; ++k) out += '0';
  for (; i < length; ++i) out += digits[(size_t)i];
  if (point >= length) out += '0';
  if (exp_mode) {
    out += 'e';
    out += exp > 0 ? '+' : '-';
    out += std::to_string(std::abs(exp - 1));
  }
  return out;
}

namespace {

// Elements of `a` (any dtype) as text, in buffer order.
std::vector<std::string> element_texts(const NArray &a, bool yaml) {
  const int64_t n = a.size();
  std::vector<std::string> t((size_t)n);
  if (n == 0) return t;
  const DType dt = a.dtype();
  std::vector<unsigned char> h((size_t)n * dtype_size(dt));
  a.to_host(h.data(), h.size());
  auto float_text = [&](double v, bool single) {
    if (yaml && std::isinf(v)) return std::string(v > 0 ? ".inf" : "-.inf");
    if (yaml && std::isnan(v)) return std::string(".nan");
    return float_to_s(v, single);
  };
  for (int64_t i = 0; i < n; ++i) {
    switch (dt) {
      case DType::I32: {
        int32_t v;
        std::memcpy(&v, &h[(size_t)i * 4], 4);
        t[(size_t)i] = std::to_string(v);
        break;
      }
      case DType::F32: {
        float v;
        std::memcpy(&v, &h[(size_t)i * 4], 4);
        t[(size_t)i] = float_text(v, true);
        break;
      }
      case DType::F64: {
        double v;
        std::memcpy(&v, &h[(size_t)i * 8], 8);
        t[(size_t)i] = float_text(v, false);
        break;
      }
      case DType::BF16: {
        uint16_t b;
        std::memcpy(&b, &h[(size_t)i * 2], 2);
        const uint32_t u = (uint32_t)b << 16;
        float v;
        std::memcpy(&v, &u, 4);
        t[(size_t)i] = float_text(v, true);
        break;
      }
      case DType::Bool:
        t[(size_t)i] = h[(size_t)i] ? "true" : "false";
        break;
    }
  }
  return t;
}

// Builds an NArray of `dtype` from element texts (already validated).
NArray build(const Shape &shape, DType dtype, const std::vector<std::string> &el, const char *fmt) {
  const size_t n = el.size();
  std::vector<unsigned char> h(n * dtype_size(dtype));
  for (size_t i = 0; i < n; ++i) {
    const std::string &s = el[i];
    auto bad = [&]() { throw ParseError(std::string("Could not read NArray from ") + fmt + ": invalid " +
                                        dtype_name(dtype) + " element '" + s + "'"); };
    switch (dtype) {
      case DType::I32: {
        int64_t v = 0;
        auto r = std::from_chars(s.data(), s.data() + s.size(), v);
        if (r.ec != std::errc() || r.ptr != s.data() + s.size() || v < INT32_MIN || v > INT32_MAX) bad();
        const int32_t w = (int32_t)v;
        std::memcpy(&h[i * 4], &w, 4);
        break;
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
          if (r.ec != std::errc() || r.ptr != s.data() + s.size()) bad();
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
        break;
      }
      case DType::Bool:
        if (s == "true") h[i] = 1;
        else if (s == "false") h[i] = 0;
        else bad();
        break;
    }
  }
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
    expect('[');
    if (eat(']')) return r;
    do {
      ws();
      if (i < s.size() && (s[i] == '[' || s[i] == '{' || s[i] == '"')) fail("expected a number or boolean");
      r.push_back(atom());
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
std::string to_json(const NArray &a) {
  std::string out = "{\"shape\":[";
  for (size_t i = 0; i < a.shape().size(); ++i) out += (i ? "," : "") + std::to_string(a.shape()[i]);
  out += "],\"elements\":[";
  const std::vector<std::string> el = element_texts(a, false);
  for (size_t i = 0; i < el.size(); ++i) {
    if (el[i] == "NaN" || el[i] == "Infinity" || el[i] == "-Infinity")
      throw ParseError(el[i] + " not allowed in JSON");
    if (i) out += ',';
    out += el[i];
  }
  return out + "]}";
}

// NArray(T).from_json / NArray.new(pull : JSON::PullParser) (n_array.cr:820-850)
NArray narray_from_json(const std::string &text, DType dtype) {
  Json j(text);
  bool found_shape = false, found_elements = false;
  std::vector<std::string> shape_t, el;
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
        el = j.array();
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
  return build(shape, dtype, el, "JSON");
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
  return build(shape, dtype, el, "YAML");
}

}  // namespace phase
