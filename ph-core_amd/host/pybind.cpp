// Python binding of the host-side Phase API (phase.hpp) for tests and the
// benchmark.  Region literals are written as Crystal source strings
// ("1..3..-1", "..-2..", "0...5") or Python ints (an Int index).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "phase.hpp"

namespace py = pybind11;
using namespace phase;

namespace {

Scalar to_scalar(const py::handle &o) {
  if (py::isinstance<Scalar>(o)) return o.cast<Scalar>();
  if (py::isinstance<py::bool_>(o)) return Scalar::boolean(o.cast<bool>());
  if (py::isinstance<py::int_>(o)) return Scalar::i32(o.cast<int64_t>());
  if (py::isinstance<py::float_>(o)) return Scalar::f64(o.cast<double>());
  throw ArgumentError("not a scalar: " + std::string(py::str(o)));
}

py::object from_scalar(const Scalar &s) {
  switch (s.dt) {
    case DType::I32: return py::int_(s.i);
    case DType::Bool: return py::bool_(s.i != 0);
    default: return py::float_(s.f);
  }
}

bool is_scalar(const py::handle &o) {
  return py::isinstance<Scalar>(o) || py::isinstance<py::bool_>(o) || py::isinstance<py::int_>(o) ||
         py::isinstance<py::float_>(o);
}

RangeLiteral to_lit(const py::handle &o) {
  if (py::isinstance<RangeLiteral>(o)) return o.cast<RangeLiteral>();
  if (py::isinstance<py::str>(o)) return RangeLiteral::parse(o.cast<std::string>());
  if (py::isinstance<py::int_>(o) && !py::isinstance<py::bool_>(o)) return RangeLiteral::of_index(o.cast<int64_t>());
  throw ArgumentError("not a range literal: " + std::string(py::str(o)));
}

RegionLiteral to_region(const py::handle &o) {
  RegionLiteral r;
  if (py::isinstance<py::tuple>(o) || py::isinstance<py::list>(o)) {
    for (auto item : o) r.push_back(to_lit(item));
  } else {
    r.push_back(to_lit(o));
  }
  return r;
}

RegionLiteral args_region(const py::args &args) {
  if (args.size() == 1 && (py::isinstance<py::tuple>(args[0]) || py::isinstance<py::list>(args[0])))
    return to_region(args[0]);
  RegionLiteral r;
  for (auto a : args) r.push_back(to_lit(a));
  return r;
}

std::vector<int64_t> args_ints(const py::args &args) {
  if (args.size() == 1 && (py::isinstance<py::tuple>(args[0]) || py::isinstance<py::list>(args[0])))
    return args[0].cast<std::vector<int64_t>>();
  std::vector<int64_t> v;
  for (auto a : args) v.push_back(a.cast<int64_t>());
  return v;
}

bool is_operand(const py::handle &o) {
  return py::isinstance<NArray>(o) || py::isinstance<View>(o);
}

Operand to_operand(const py::handle &o) {
  if (py::isinstance<ProcView>(o)) return o.cast<const ProcView &>().operand();
  if (py::isinstance<View>(o)) return o.cast<const View &>().operand();
  if (py::isinstance<NArray>(o)) return o.cast<const NArray &>().operand();
  throw ArgumentError("not a MultiIndexable: " + std::string(py::str(o)));
}

DType value_dtype_of(const py::handle &o) { return to_operand(o).value_dtype(); }

py::dtype np_dtype(DType d) {
  switch (d) {
    case DType::I32: return py::dtype::of<int32_t>();
    case DType::F32: return py::dtype::of<float>();
    case DType::F64: return py::dtype::of<double>();
    case DType::BF16: return py::dtype::of<uint16_t>();
    case DType::Bool: return py::dtype::of<bool>();
  }
  return py::dtype::of<double>();
}

DType dtype_of_np(const py::dtype &d) {
  if (d.is(py::dtype::of<int32_t>())) return DType::I32;
  if (d.is(py::dtype::of<float>())) return DType::F32;
  if (d.is(py::dtype::of<double>())) return DType::F64;
  if (d.is(py::dtype::of<bool>())) return DType::Bool;
  if (d.is(py::dtype::of<uint16_t>())) return DType::BF16;
  throw ArgumentError("unsupported numpy dtype " + std::string(py::str(d)));
}

NArray from_numpy(const py::array &arr, std::optional<DType> dt) {
  DType d = dt ? *dt : dtype_of_np(arr.dtype());
  py::array c = py::array::ensure(arr, py::array::c_style);
  if (!c.dtype().is(np_dtype(d))) c = c.attr("astype")(np_dtype(d));
  c = py::array::ensure(c, py::array::c_style);
  Shape shape(c.shape(), c.shape() + c.ndim());
  return NArray::of_buffer(shape, d, c.data(), (size_t)c.nbytes());
}

py::array to_numpy(const NArray &a) {
  py::array out(np_dtype(a.dtype()), a.shape());
  if (a.size()) a.to_host(out.mutable_data(), (size_t)out.nbytes());
  return out;
}

// Result of a binary operator between an operand and a Python object.
py::object binop(const std::string &op, const py::handle &self, const py::handle &other, bool reflected) {
  Operand a = to_operand(self);
  if (is_scalar(other)) {
    Scalar s = to_scalar(other);
    return py::cast(reflected ? elementwise(op, s, a) : elementwise(op, a, s));
  }
  if (is_operand(other)) {
    Operand b = to_operand(other);
    return py::cast(reflected ? elementwise(op, b, a) : elementwise(op, a, b));
  }
  return py::reinterpret_borrow<py::object>(Py_NotImplemented);
}

Expr as_expr(const py::handle &o) {
  if (py::isinstance<Expr>(o)) return o.cast<Expr>();
  return Expr::constant(to_scalar(o));
}

// Calls a map_with_coord block with (element, coordinate) or (element,
// coordinate, index), as many as it declares: a Crystal block ignores yielded
// values it has no parameter for.  The coordinate is a list of Int32
// expressions, one per axis.
Expr call_coord_block(const py::function &fn, const Expr &x, const std::vector<Expr> &c, const Expr &i) {
  py::list coord;
  for (const Expr &e : c) coord.append(py::cast(e));
  size_t nparams = 2;
  try {
    nparams = py::len(py::module_::import("inspect").attr("signature")(fn).attr("parameters"));
  } catch (const py::error_already_set &) {
  }
  return nparams >= 3 ? as_expr(fn(x, coord, i)) : as_expr(fn(x, coord));
}

template <typename C>
void def_operators(C &cls) {
  struct Op {
    const char *py;
    const char *cr;
    bool reflected;
  };
  static const Op ops[] = {
      {"__add__", "+", false},   {"__radd__", "+", true},     {"__sub__", "-", false},
      {"__rsub__", "-", true},   {"__mul__", "*", false},     {"__rmul__", "*", true},
      {"__truediv__", "/", false}, {"__rtruediv__", "/", true}, {"__floordiv__", "//", false},
      {"__rfloordiv__", "//", true}, {"__mod__", "%", false},  {"__rmod__", "%", true},
      {"__pow__", "**", false},  {"__rpow__", "**", true},    {"__and__", "&", false},
      {"__rand__", "&", true},   {"__or__", "|", false},      {"__ror__", "|", true},
      {"__xor__", "^", false},   {"__rxor__", "^", true},     {"__gt__", ">", false},
      {"__lt__", "<", false},    {"__ge__", ">=", false},     {"__le__", "<=", false},
      {"wadd", "&+", false},     {"wsub", "&-", false},       {"wmul", "&*", false},
      {"wpow", "&**", false},    {"cmp", "<=>", false},
  };
  for (const Op &o : ops) {
    std::string cr = o.cr;
    bool refl = o.reflected;
    cls.def(o.py, [cr, refl](py::object self, py::object other) { return binop(cr, self, other, refl); },
            py::is_operator());
  }
  cls.def("__neg__", [](py::object self) { return elementwise_unary("-", to_operand(self)); });
  cls.def("__pos__", [](py::object self) { return elementwise_unary("+", to_operand(self)); });
  cls.def("__invert__", [](py::object self) { return elementwise_unary("~", to_operand(self)); });
  cls.def("__eq__", [](py::object self, py::object other) {
    if (!is_operand(other)) return false;
    return equals(to_operand(self), to_operand(other));
  });
  cls.def("__ne__", [](py::object self, py::object other) {
    if (!is_operand(other)) return true;
    return !equals(to_operand(self), to_operand(other));
  });
  cls.def("eq", [](py::object self, py::object other) -> NArray {
    if (is_operand(other)) return eq(to_operand(self), to_operand(other));
    Scalar s = to_scalar(other);
    return map(to_operand(self), [s](const Expr &x) { return Expr::binary("==", x, Expr::constant(s)); });
  });
  cls.def("equals", [](py::object self, py::object other) { return equals(to_operand(self), to_operand(other)); });
  cls.def("map", [](py::object self, py::function fn) {
    return map(to_operand(self), [&](const Expr &x) { return as_expr(fn(x)); });
  });
  cls.def("map_with_coord", [](py::object self, py::function fn) {
    return map_with_coord(to_operand(self), [&](const Expr &x, const std::vector<Expr> &c, const Expr &i) {
      return call_coord_block(fn, x, c, i);
    });
  });
  cls.def("map_with", [](py::object self, py::args others, py::kwargs kw) {
    py::function fn = kw["block"].cast<py::function>();
    std::vector<Operand> ops{to_operand(self)};
    for (auto o : others) ops.push_back(to_operand(o));
    return map_with(ops, [&](const std::vector<Expr> &x) {
      py::list l;
      for (const Expr &e : x) l.append(py::cast(e));
      return as_expr(fn(*l));
    });
  });
  cls.def("sum", [](py::object self, const std::string &acc) {
        return from_scalar(sum(to_operand(self), acc == "f64" ? Acc::F64 : Acc::F32C));
      }, py::arg("acc") = "f32c");
  cls.def("max", [](py::object self) { return from_scalar(max(to_operand(self))); });
  cls.def("slices_sum", [](py::object self, int axis, const std::string &acc) {
        return slices_sum(to_operand(self), axis, acc == "f64" ? Acc::F64 : Acc::F32C);
      }, py::arg("axis") = 0, py::arg("acc") = "f32c");
  cls.def("slices_max", [](py::object self, int axis) { return slices_max(to_operand(self), axis); },
          py::arg("axis") = 0);
  cls.def("reduce2d_all", [](py::object self, const std::string &acc) {
        Reduce2D r = reduce2d_all(to_operand(self), acc == "f64" ? Acc::F64 : Acc::F32C);
        py::dict d;
        d["full_sum"] = r.full_sum;
        d["full_max"] = r.full_max;
        d["row_sum"] = r.row_sum;
        d["row_max"] = r.row_max;
        d["col_sum"] = r.col_sum;
        d["col_max"] = r.col_max;
        return d;
      }, py::arg("acc") = "f32c");
  cls.def("reduce2d_all_device", [](py::object self, const std::string &acc) {
        return reduce2d_all_device(to_operand(self), acc == "f64" ? Acc::F64 : Acc::F32C);
      }, py::arg("acc") = "f32c");
  cls.def_property_readonly("value_dtype", [](py::object self) { return value_dtype_of(self); });
}

}  // namespace

PYBIND11_MODULE(_phase, m) {
  m.doc() = "MI355X backend for ph-core's NArray data-parallel path (host API)";

  auto base = py::register_exception<PhaseError>(m, "PhaseError");
  py::register_exception<IndexError>(m, "IndexError", PyExc_IndexError);
  auto shape_err = py::register_exception<ShapeError>(m, "ShapeError", base.ptr());
  py::register_exception<DimensionError>(m, "DimensionError", shape_err.ptr());
  py::register_exception<ArgumentError>(m, "ArgumentError", PyExc_ValueError);
  py::register_exception<OverflowError>(m, "OverflowError", PyExc_OverflowError);
  py::register_exception<DivisionByZeroError>(m, "DivisionByZeroError", PyExc_ZeroDivisionError);
  py::register_exception<EmptyError>(m, "EmptyError", base.ptr());
  py::register_exception<DeviceError>(m, "DeviceError", base.ptr());
  py::register_exception<ParseError>(m, "ParseError", PyExc_ValueError);

  py::enum_<DType>(m, "DType")
      .value("I32", DType::I32)
      .value("F32", DType::F32)
      .value("F64", DType::F64)
      .value("BF16", DType::BF16)
      .value("Bool", DType::Bool);

  py::class_<Scalar>(m, "Scalar")
      .def_readonly("dtype", &Scalar::dt)
      .def("value", [](const Scalar &s) { return from_scalar(s); })
      .def("__repr__", &Scalar::to_s);
  m.def("f32", [](double v) { return Scalar::f32(v); }, "a Float32 literal");
  m.def("f64", [](double v) { return Scalar::f64(v); }, "a Float64 literal");
  m.def("i32", [](int64_t v) { return Scalar::i32(v); }, "an Int32 literal");

  m.def("init", &phase::init, py::arg("device") = 0);
  m.def("device_available", &phase::device_available);
  m.def("jit_compiles", []() { return ph_jit_compiles(); });
  m.def("launch_count", []() { return ph_launch_count(); }, "kernel launches issued so far");
  m.def("sync", []() {
    if (ph_sync() != 0) throw std::runtime_error(ph_last_error());
  }, "waits for the library stream");

  // --- range syntax / utils
  py::class_<RangeLiteral>(m, "RangeLiteral")
      .def_static("parse", &RangeLiteral::parse)
      .def_readonly("is_index", &RangeLiteral::is_index)
      .def_readonly("index", &RangeLiteral::index)
      .def_property_readonly("first", [](const RangeLiteral &r) { return r.first; })
      .def_property_readonly("last", [](const RangeLiteral &r) { return r.last; })
      .def_property_readonly("step", [](const RangeLiteral &r) { return r.step; })
      .def_readonly("exclusive", &RangeLiteral::exclusive)
      .def("__repr__", &RangeLiteral::to_s);
  py::module_ rs = m.def_submodule("RangeSyntax");
  rs.def("get_size", &RangeSyntax::get_size);
  rs.def("infer_range", [](py::object lit, int64_t bound) {
    CanonicalRange c = RangeSyntax::infer_range(to_lit(lit), bound);
    return py::make_tuple(c.first, c.step, c.last, c.size);
  });
  rs.def("canonicalize_range", [](py::object lit, int64_t bound) {
    CanonicalRange c = RangeSyntax::canonicalize_range(to_lit(lit), bound);
    return py::make_tuple(c.first, c.step, c.last, c.size);
  });
  py::module_ cu = m.def_submodule("CoordUtil");
  cu.def("has_index", &CoordUtil::has_index);
  cu.def("canonicalize_index", &CoordUtil::canonicalize_index);
  cu.def("canonicalize_coord", &CoordUtil::canonicalize_coord);
  py::module_ su = m.def_submodule("ShapeUtil");
  su.def("compatible_shapes", &ShapeUtil::compatible_shapes);
  su.def("shape_to_size", &ShapeUtil::shape_to_size);

  // --- IndexRegion
  py::class_<IndexRegion>(m, "IndexRegion")
      .def(py::init([](py::object lit, const Shape &bound, bool drop) {
             return IndexRegion::from_literal(to_region(lit), bound, drop);
           }),
           py::arg("region_literal"), py::arg("bound_shape"), py::arg("drop") = true)
      .def_static("absolute", [](py::object lit, bool drop) { return IndexRegion::absolute(to_region(lit), drop); },
                  py::arg("region_literal"), py::arg("drop") = true)
      .def_static("trimmed",
                  [](py::object lit, std::optional<Shape> bound, bool drop, const Shape &trim_to) {
                    return IndexRegion::trimmed(to_region(lit), bound ? &*bound : nullptr, drop, trim_to);
                  },
                  py::arg("region_literal"), py::arg("bound_shape"), py::arg("drop"), py::arg("trim_to"))
      .def_static("cover",
                  [](const Shape &s, bool drop, std::optional<std::vector<bool>> degen) {
                    return IndexRegion::cover(s, drop, degen ? &*degen : nullptr);
                  },
                  py::arg("shape"), py::arg("drop") = true, py::arg("degeneracy") = py::none())
      .def_static("fitted", &IndexRegion::fitted_copy)
      .def_readonly("first", &IndexRegion::first)
      .def_readonly("step", &IndexRegion::step)
      .def_readonly("last", &IndexRegion::last)
      .def_readonly("proper_shape", &IndexRegion::proper_shape)
      .def_readonly("degeneracy", &IndexRegion::degeneracy)
      .def_readonly("drop", &IndexRegion::drop)
      .def_property_readonly("shape", &IndexRegion::shape)
      .def_property_readonly("size", &IndexRegion::size)
      .def_property_readonly("dimensions", &IndexRegion::dimensions)
      .def_property_readonly("proper_dimensions", &IndexRegion::proper_dimensions)
      .def("trim", [](IndexRegion &r, const Shape &b) -> IndexRegion & { return r.trim(b); },
           py::return_value_policy::reference_internal)
      .def("reverse", [](IndexRegion &r) -> IndexRegion & { return r.reverse(); },
           py::return_value_policy::reference_internal)
      .def("translate", [](IndexRegion &r, const std::vector<int64_t> &o) -> IndexRegion & { return r.translate(o); },
           py::return_value_policy::reference_internal)
      .def("unsafe_translate",
           [](IndexRegion &r, const std::vector<int64_t> &o) -> IndexRegion & { return r.unsafe_translate(o); },
           py::return_value_policy::reference_internal)
      .def("includes", &IndexRegion::includes)
      .def("fits_in", &IndexRegion::fits_in)
      .def("local_to_absolute", &IndexRegion::local_to_absolute)
      .def("local_to_absolute_unsafe", &IndexRegion::local_to_absolute_unsafe)
      .def("absolute_to_local", &IndexRegion::absolute_to_local)
      .def("absolute_to_local_unsafe", &IndexRegion::absolute_to_local_unsafe)
      .def("copy", [](const IndexRegion &r) { return r; })
      .def("__eq__", &IndexRegion::operator==)
      .def("__repr__", &IndexRegion::to_s);

  // --- expressions (ProcView blocks)
  py::class_<Expr> ex(m, "Expr");
  ex.def_property_readonly("dtype", &Expr::dtype)
      .def("__repr__", &Expr::to_s)
      .def("abs", [](const Expr &e) { return Expr::unary("abs", e); })
      .def("to_f32", [](const Expr &e) { return Expr::unary("to_f32", e); })
      .def("to_f64", [](const Expr &e) { return Expr::unary("to_f64", e); })
      .def("to_i32", [](const Expr &e) { return Expr::unary("to_i32", e); })
      .def("__neg__", [](const Expr &e) { return Expr::unary("-", e); })
      .def("__pos__", [](const Expr &e) { return Expr::unary("+", e); })
      .def("__invert__", [](const Expr &e) { return Expr::unary("~", e); });
  {
    struct Op {
      const char *py;
      const char *cr;
      bool reflected;
    };
    static const Op ops[] = {
        {"__add__", "+", false},     {"__radd__", "+", true},       {"__sub__", "-", false},
        {"__rsub__", "-", true},     {"__mul__", "*", false},       {"__rmul__", "*", true},
        {"__truediv__", "/", false}, {"__rtruediv__", "/", true},   {"__floordiv__", "//", false},
        {"__rfloordiv__", "//", true}, {"__mod__", "%", false},     {"__rmod__", "%", true},
        {"__pow__", "**", false},    {"__rpow__", "**", true},      {"__and__", "&", false},
        {"__rand__", "&", true},     {"__or__", "|", false},        {"__ror__", "|", true},
        {"__xor__", "^", false},     {"__rxor__", "^", true},       {"__gt__", ">", false},
        {"__lt__", "<", false},      {"__ge__", ">=", false},       {"__le__", "<=", false},
        {"__eq__", "==", false},     {"__ne__", "!=", false},       {"wadd", "&+", false},
        {"wsub", "&-", false},       {"wmul", "&*", false},         {"wpow", "&**", false},
        {"cmp", "<=>", false},
    };
    for (const Op &o : ops) {
      std::string cr = o.cr;
      bool refl = o.reflected;
      ex.def(o.py, [cr, refl](const Expr &a, py::object b) {
        Expr e = as_expr(b);
        return refl ? Expr::binary(cr, e, a) : Expr::binary(cr, a, e);
      }, py::is_operator());
    }
    ex.attr("__hash__") = py::none();
  }
  m.def("sqrt", [](py::object e) { return Expr::unary("sqrt", as_expr(e)); });
  m.def("select", [](py::object c, py::object a, py::object b) {
    return Expr::select(as_expr(c), as_expr(a), as_expr(b));
  });
  m.def("placeholder", [](DType d) { return Expr::input(0, d); }, "the element of a one-input block");
  m.def("input", [](int k, DType d) { return Expr::input(k, d); }, "element of operand k");
  m.def("element_index", []() { return Expr::index(); }, "the element's lexicographic index (Int32)");
  m.def("compile_check", [](const Expr &e, const std::vector<DType> &in_dt, int64_t rows, int64_t cols, int kind) {
    ph_prog p = e.compile((int)in_dt.size());
    std::vector<int32_t> d;
    for (DType t : in_dt) d.push_back((int32_t)t);
    check(ph_compile_check(&p, (int32_t)d.size(), d.data(), rows, cols, kind), "compile_check");
  }, "generate and compile (no launch) the fused kernel for an expression");
  m.def("program", [](const Expr &e, int n_in) {
    ph_prog p = e.compile(n_in);
    py::list l;
    for (int i = 0; i < p.n_insn; ++i) {
      const ph_insn &in = p.insn[i];
      l.append(py::make_tuple((int)in.op, (int)in.dt, (int)in.dst, (int)in.a, (int)in.b, (int)in.c,
                              in.dt == PH_F32 || in.dt == PH_F64 ? py::cast(in.imm.f) : py::cast(in.imm.i)));
    }
    return py::make_tuple(l, p.out_reg, p.out_dtype);
  }, "the lowered program (for inspection)");

  // --- NArray
  py::class_<NArray> na(m, "NArray");
  na.def_static("fill", [](const Shape &s, py::object v, std::optional<DType> dt) {
        return NArray::fill(s, to_scalar(v), dt);
      }, py::arg("shape"), py::arg("value"), py::arg("dtype") = py::none())
      .def_static("of_buffer", [](const Shape &shape, py::array buf, std::optional<DType> dt) {
        NArray a = from_numpy(buf.attr("reshape")(-1), dt);
        return a.reshape(shape);
      }, py::arg("shape"), py::arg("buffer"), py::arg("dtype") = py::none())
      .def_static("from_numpy", &from_numpy, py::arg("array"), py::arg("dtype") = py::none())
      .def_static("random",
                  [](const Shape &s, DType d, uint64_t seed, uint64_t offset, const std::string &dist, double a,
                     double b) {
                    int k = dist == "uniform" ? 0 : dist == "normal" ? 1 : dist == "int31" ? 2 : -1;
                    if (k < 0) throw ArgumentError("dist must be uniform, normal or int31");
                    return NArray::random(s, d, seed, offset, k, a, b);
                  },
                  py::arg("shape"), py::arg("dtype"), py::arg("seed"), py::arg("offset") = 0,
                  py::arg("dist") = "uniform", py::arg("a") = 0.0, py::arg("b") = 1.0)
      .def_static("empty", &NArray::empty)
      .def("fill_random",
           [](NArray &self, uint64_t seed, const std::string &dist, double a, double b) {
             int k = dist == "uniform" ? 0 : dist == "normal" ? 1 : dist == "int31" ? 2 : -1;
             if (k < 0) throw ArgumentError("dist must be uniform, normal or int31");
             check(ph_fill_random(self.device_ptr(), (int32_t)self.dtype(), self.size(), seed, 0, k, a, b),
                   "random fill");
           },
           py::arg("seed"), py::arg("dist") = "uniform", py::arg("a") = 0.0, py::arg("b") = 1.0,
           "overwrites every element with draws of stream `seed` (as NArray.random)")
      .def_property_readonly("shape", &NArray::shape)
      .def_property_readonly("size", &NArray::size)
      .def_property_readonly("dimensions", &NArray::dimensions)
      .def_property_readonly("dtype", &NArray::dtype)
      .def("is_empty", &NArray::empty_)
      .def("is_scalar", &NArray::scalar_)
      .def("to_numpy", &to_numpy)
      .def("get", [](const NArray &a, py::args c) { return from_scalar(a.get(args_ints(c))); })
      .def("materialize", [](NArray &a) -> NArray & {
        a.storage->dev_read();  // runs a pending operator result (or fill) now
        return a;
      }, "issues the array's pending work on the device (an operator result left pending, a fill)")
      .def("set_element", [](NArray &a, const Coord &c, py::object v) { a.set_element(c, to_scalar(v)); })
      .def("__getitem__", [](const NArray &a, py::object key) { return a.get_chunk(to_region(key)); })
      .def("get_chunk", [](const NArray &a, py::object lit, bool drop) { return a.get_chunk(to_region(lit), drop); },
           py::arg("region_literal"), py::arg("drop") = true)
      .def("get_chunk_region", [](const NArray &a, const IndexRegion &r) { return a.get_chunk(r); })
      .def("unsafe_fetch_chunk", &NArray::unsafe_fetch_chunk)
      .def("get_available", [](const NArray &a, py::object lit, bool drop) { return a.get_available(to_region(lit), drop); },
           py::arg("region_literal"), py::arg("drop") = true)
      .def("__setitem__", [](NArray &a, py::object key, py::object value) {
        if (py::isinstance<NArray>(key) || py::isinstance<View>(key)) {  // mask
          Operand mask = to_operand(key);
          if (is_scalar(value)) a.set_mask(mask, to_scalar(value));
          else a.set_mask(mask, to_operand(value));
          return;
        }
        RegionLiteral lit = to_region(key);
        if (is_scalar(value)) a.set_chunk(lit, to_scalar(value));
        else a.set_chunk(lit, to_operand(value));
      })
      .def("unsafe_set_chunk", [](NArray &a, const IndexRegion &r, py::object value) {
        if (is_scalar(value)) a.unsafe_set_chunk(r, to_scalar(value));
        else a.unsafe_set_chunk(r, to_operand(value));
      })
      .def("view", [](const NArray &a, py::args lit) { return lit.size() ? a.view(args_region(lit)) : a.view(); })
      .def("mutable_view",
           [](const NArray &a, py::args lit) { return lit.size() ? a.mutable_view(args_region(lit)) : a.mutable_view(); })
      .def("process", [](const NArray &a, py::function fn) {
        return a.process(as_expr(fn(Expr::input(0, a.dtype()))));
      })
      .def("tile", [](const NArray &a, py::args c) { return a.tile(args_ints(c)); })
      .def("reshape", [](const NArray &a, py::args s) { return a.reshape(args_ints(s)); })
      .def("flatten", &NArray::flatten)
      .def("permute", [](const NArray &a, py::args o) {
        if (!o.size()) return a.permute(nullptr);
        std::vector<int64_t> v = args_ints(o);
        return a.permute(&v);
      })
      .def("reverse", &NArray::reverse)
      .def("clone", &NArray::clone)
      .def_property_readonly("device_ptr", [](const NArray &a) { return (uintptr_t)a.device_ptr(); },
                             "address of the device buffer (for C-ABI calls)")
      .def("to_json", [](const NArray &a) { return to_json(a); })
      .def("to_yaml", [](const NArray &a) { return to_yaml(a); })
      .def_static("from_json", &narray_from_json, py::arg("text"), py::arg("dtype"))
      .def_static("from_yaml", &narray_from_yaml, py::arg("text"), py::arg("dtype"))
      .def_static("concatenate_all",
                  [](py::args narrs, int64_t axis) {
                    std::vector<Operand> v;
                    for (auto o : narrs) v.push_back(to_operand(o));
                    return NArray::concatenate(v, axis);
                  },
                  "NArray.concatenate(*narrs, axis: 0)", py::kw_only(), py::arg("axis") = 0)
      .def("concatenate",
           [](const NArray &a, py::args others, int64_t axis) {
             std::vector<Operand> v;
             for (auto o : others) v.push_back(to_operand(o));
             return a.concatenate_with(v, axis);
           },
           py::kw_only(), py::arg("axis") = 0)
      .def("concatenate_",
           [](NArray &a, py::args others, int64_t axis) -> NArray & {
             std::vector<Operand> v;
             for (auto o : others) v.push_back(to_operand(o));
             return a.concatenate_inplace(v, axis);
           },
           py::kw_only(), py::arg("axis") = 0, py::return_value_policy::reference_internal)
      .def("push",
           [](NArray &a, py::args others) -> NArray & {
             std::vector<Operand> v;
             for (auto o : others) v.push_back(to_operand(o));
             return a.push(v);
           },
           py::return_value_policy::reference_internal)
      .def("to_narr", &NArray::clone)
      .def("slices", &NArray::slices, py::arg("axis") = 0)
      .def("map_", [](NArray &a, py::function fn) -> NArray & {
        return a.map_inplace([&](const Expr &x) { return as_expr(fn(x)); });
      }, py::return_value_policy::reference_internal)
      .def("map_with_index", [](const NArray &a, py::function fn) {
        return a.map_with_index([&](const Expr &x, const Expr &i) { return as_expr(fn(x, i)); });
      })
      .def("map_with_index_", [](NArray &a, py::function fn) -> NArray & {
        return a.map_with_index_inplace([&](const Expr &x, const Expr &i) { return as_expr(fn(x, i)); });
      }, py::return_value_policy::reference_internal)
      .def("map_with_coord_", [](NArray &a, py::function fn) -> NArray & {
        return a.map_with_coord_inplace([&](const Expr &x, const std::vector<Expr> &c, const Expr &i) {
          return call_coord_block(fn, x, c, i);
        });
      }, py::return_value_policy::reference_internal)
      .def("__repr__", [](const NArray &a) {
        return std::string("<NArray ") + dtype_name(a.dtype()) + " " + shape_to_s(a.shape()) + ">";
      });
  def_operators(na);

  // --- views
  py::class_<View> vw(m, "View");
  vw.def_property_readonly("shape", &View::shape)
      .def_property_readonly("size", &View::size)
      .def_property_readonly("dimensions", &View::dimensions)
      .def_property_readonly("dtype", &View::dtype)
      .def("to_narr", &View::to_narr)
      .def("to_numpy", [](const View &v) { return to_numpy(v.to_narr()); })
      .def("view", [](const View &v, py::args lit) { return lit.size() ? v.view(args_region(lit)) : v; })
      .def("__getitem__", [](const View &v, py::object key) { return v.view(to_region(key)); })
      .def("reshape", [](const View &v, py::args s) { return v.reshape(args_ints(s)); })
      .def("permute", [](const View &v, py::args o) {
        if (!o.size()) return v.permute(nullptr);
        std::vector<int64_t> p = args_ints(o);
        return v.permute(&p);
      })
      .def("reverse", &View::reverse)
      .def("tile", [](const View &v, py::args c) {
        View r = v;
        r.tile_(args_ints(c));
        return r;
      })
      .def("process", [](const View &v, py::function fn) {
        return v.process(as_expr(fn(Expr::input(0, v.dtype()))));
      })
      .def("apply", &View::apply)
      .def("get", [](const View &v, py::args c) { return from_scalar(v.get(args_ints(c))); })
      .def("descriptor", [](const View &v) {
        const ph_view_desc &d = v.desc();
        py::dict r;
        r["tail"] = v.tail_;  // transforms past the descriptor's stages
        r["rank"] = d.view.rank;
        r["offset"] = d.view.offset;
        r["shape"] = std::vector<int64_t>(d.view.shape, d.view.shape + d.view.rank);
        r["stride"] = std::vector<int64_t>(d.view.stride, d.view.stride + d.view.rank);
        r["nstage"] = d.nstage;
        return r;
      });
  def_operators(vw);

  py::class_<MutableView, View>(m, "MutableView")
      .def("mutable_view", [](const MutableView &v, py::args lit) {
        return lit.size() ? v.mutable_view(args_region(lit)) : v;
      })
      .def("__setitem__", [](MutableView &v, py::object key, py::object value) {
        RegionLiteral lit = to_region(key);
        if (is_scalar(value)) v.set_chunk(lit, to_scalar(value));
        else v.set_chunk(lit, to_operand(value));
      })
      .def("unsafe_set_chunk", [](MutableView &v, const IndexRegion &r, py::object value) {
        v.unsafe_set_chunk(r, to_operand(value));
      })
      // View#reshape / permute / reverse clone self: on a MutableView they stay mutable
      .def("reshape", [](const MutableView &v, py::args s) {
        MutableView m = v;
        m.reshape_(args_ints(s));
        return m;
      })
      .def("permute", [](const MutableView &v, py::args o) {
        MutableView m = v;
        if (!o.size()) {
          m.permute_(nullptr);
        } else {
          std::vector<int64_t> p = args_ints(o);
          m.permute_(&p);
        }
        return m;
      })
      .def("reverse", [](const MutableView &v) {
        MutableView m = v;
        m.reverse_();
        return m;
      })
      .def("set_element", [](MutableView &v, const Coord &c, py::object val) { v.set_element(c, to_scalar(val)); });

  py::class_<ProcView, View>(m, "ProcView")
      .def("to_narr", &ProcView::to_narr)
      .def("to_numpy", [](const ProcView &v) { return to_numpy(v.to_narr()); })
      .def("process", [](const ProcView &v, py::function fn) {
        return v.process(as_expr(fn(Expr::input(0, v.body.dtype()))));
      })
      .def("get", [](const ProcView &v, py::args c) { return from_scalar(v.get(args_ints(c))); })
      .def_property_readonly("body", [](const ProcView &v) { return v.body; });

  m.def("float_to_s", &phase::float_to_s, py::arg("value"), py::arg("single") = false);
  m.def("heat2d", &phase::heat2d, py::arg("u"), py::arg("r"), py::arg("steps"), py::arg("every") = 0);
  m.def("mirror_stats", []() {
    const phase::MirrorStats st = phase::mirror_stats();
    py::dict d;
    d["pages_fetched"] = st.pages_fetched;
    d["bytes_fetched"] = st.bytes_fetched;
    d["runs_written"] = st.runs_written;
    d["bytes_written"] = st.bytes_written;
    return d;
  }, "Host-mirror traffic since the process started (pages fetched, dirty runs written back, bytes).");

  py::class_<Batch>(m, "Batch")
      .def(py::init([](const std::vector<Shape> &shapes, const std::vector<float> &scalars) {
             return std::make_unique<Batch>(shapes, scalars);
           }),
           py::arg("shapes"), py::arg("scalars"))
      .def("__len__", &Batch::count)
      .def("x_at", &Batch::x_at)
      .def("y_at", &Batch::y_at)
      .def("z_at", &Batch::z_at)
      .def("fill_uniform", &Batch::fill_uniform, py::arg("seed"), py::arg("a") = -1.0, py::arg("b") = 1.0)
      .def("run", &Batch::run, py::arg("mode") = 0)
      .def("tune", &Batch::tune)
      .def("schedule", &Batch::schedule, py::arg("chunk"), py::arg("grid"), py::arg("dynamic") = 1)
      .def("sums", &Batch::sums)
      .def_readonly("shapes", &Batch::shapes)
      .def_readonly("offsets_x", &Batch::offsets_x)
      .def_readonly("offsets_z", &Batch::offsets_z)
      .def_property_readonly("x", [](const Batch &b) { return b.x; })
      .def_property_readonly("y", [](const Batch &b) { return b.y; })
      .def_property_readonly("z", [](const Batch &b) { return b.z; });

  // progs: callables (v, s) -> expression over the element and the scalar;
  // records: dicts with "shape", and optionally "map", "red" (program
  // indices), "ops" ("sum", "max" or both, as a string or a list), "region"
  // (range literals, as `view` takes them) and "scalar".
  py::class_<GBatch>(m, "GBatch")
      .def(py::init([](DType xdt, DType ydt, const py::list &progs, const py::list &records) {
             std::vector<Expr> ps;
             for (auto f : progs) {
               if (py::isinstance<Expr>(f)) {
                 ps.push_back(f.cast<Expr>());
                 continue;
               }
               py::object r = f(py::cast(Expr::input(0, xdt)), py::cast(Expr::input(1, xdt)));
               ps.push_back(as_expr(r));
             }
             std::vector<GBatchRecord> rs;
             for (auto item : records) {
               py::dict d = item.cast<py::dict>();
               GBatchRecord g;
               g.shape = d["shape"].cast<Shape>();
               if (d.contains("map") && !d["map"].is_none()) g.map = d["map"].cast<int>();
               if (d.contains("red") && !d["red"].is_none()) g.red = d["red"].cast<int>();
               if (d.contains("ops")) {
                 py::object o = d["ops"];
                 std::vector<std::string> names;
                 if (py::isinstance<py::str>(o)) names.push_back(o.cast<std::string>());
                 else names = o.cast<std::vector<std::string>>();
                 for (const std::string &n : names) {
                   if (n == "sum") g.red_ops |= PH_RED_SUM;
                   else if (n == "max") g.red_ops |= PH_RED_MAX;
                   else throw ArgumentError("batch reduction " + n + " (sum or max)");
                 }
               }
               if (d.contains("region") && !d["region"].is_none()) {
                 g.slice = true;
                 g.region = to_region(d["region"]);
               }
               if (d.contains("scalar")) g.scalar = d["scalar"].cast<double>();
               rs.push_back(g);
             }
             return std::make_unique<GBatch>(xdt, ydt, ps, rs);
           }),
           py::arg("x_dtype"), py::arg("y_dtype"), py::arg("progs"), py::arg("records"))
      .def("__len__", &GBatch::count)
      .def("x_at", &GBatch::x_at)
      .def("y_at", &GBatch::y_at)
      .def("z_at", &GBatch::z_at)
      .def("set_x", &GBatch::set_x)
      .def("run", &GBatch::run, py::arg("mode") = 0)
      .def("tune", &GBatch::tune)
      .def("sums", &GBatch::sums)
      .def("maxes", &GBatch::maxes)
      .def_readonly("offsets_xy", &GBatch::offsets_xy)
      .def_readonly("offsets_z", &GBatch::offsets_z)
      .def_readonly("z_shapes", &GBatch::z_shapes)
      .def_property_readonly("x", [](const GBatch &b) { return b.x; })
      .def_property_readonly("y", [](const GBatch &b) { return b.y; })
      .def_property_readonly("z", [](const GBatch &b) { return b.z; });
}
