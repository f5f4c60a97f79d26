// Fused elementwise programs: ph_map, ph_reduce, ph_reduce2d_all.
//
// A ph_prog (the compiled form of an elementwise expression or a ProcView
// block) is turned into HIP source with every operand's collapsed shape and
// strides baked in as constants, compiled once with hipRTC (jit.cpp) and
// launched on the library stream.  Each instruction becomes one typed C++
// statement, so the kernel does exactly the roundings the program asks for
// (compiled with -ffp-contract=off: no fused multiply-adds).
//
// Map kernel: `rows x cols` after collapsing; a 256-thread workgroup owns
// 256*EPT consecutive columns of one row, lane t handles columns
// tile0 + e*256 + t (coalesced for every unrolled e); all loads of a tile
// are issued before any arithmetic.
//
// Reduction kernel (2-D view rows x cols): a workgroup owns TR rows x
// 256*VEC columns; wavefront w owns columns [64w*VEC, 64(w+1)*VEC), lane l
// owns VEC consecutive columns (VEC = 4, or 8 for BFloat16 and Bool inputs).
// Per element the program's value feeds (a) the thread's column sums and
// maxima, accumulated row after row, and (b) a per-row lane value that a
// fixed cross-lane tree folds into one value per wavefront and row.  Second
// pass: per-row and per-column states are folded across tiles in tile order,
// then the full result across rows in row order.  Integer folds are in
// lexicographic order, so Int32 running-sum overflow is detected exactly as a
// sequential checked sum would; every result is deterministic.
#include <algorithm>
#include <utility>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <map>
#include <regex>
#include <string>
#include <vector>

#include "common.hpp"
#include "gen.hpp"
#include "jit.hpp"

// Program lowering shared with the general batch (gen.hpp).
namespace ph {
namespace gen {

const char *ctype(int dt) {
  switch (dt) {
    case PH_I32: return "int";
    case PH_F32: return "float";
    case PH_F64: return "double";
    case PH_BOOL: return "bool";
  }
  return "void";
}

const char *stype(int dt) {  // storage type
  switch (dt) {
    case PH_I32: return "int";
    case PH_F32: return "float";
    case PH_F64: return "double";
    case PH_BF16: return "unsigned short";
    case PH_BOOL: return "unsigned char";
  }
  return "void";
}

int compute_dt(int storage) { return storage == PH_BF16 ? PH_F32 : storage; }

std::string fmt(const char *f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  std::vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

std::string lit_i64(int64_t v) {
  if (v == INT64_MIN) return "(-9223372036854775807LL-1)";
  return fmt("%" PRId64 "LL", v);
}

std::string const_expr(int dt, const ph_insn &in) {
  switch (dt) {
    case PH_I32: {
      int32_t v = (int32_t)in.imm.i;
      if (v == INT32_MIN) return "(-2147483647-1)";
      return fmt("%d", v);
    }
    case PH_F32: return fmt("((float)%a)", (double)(float)in.imm.f);
    case PH_F64: return fmt("(%a)", in.imm.f);
    case PH_BOOL: return in.imm.i ? "true" : "false";
  }
  return "0";
}

bool is_float(int dt) { return dt == PH_F32 || dt == PH_F64; }

// Validates `p` against the input storage dtypes and checks every operand's
// dtype; fills reg dtypes on the way.  PH_OP_INDEX (an element's lexicographic
// index) is accepted only where `allow_index` says the kernel defines it.
int validate_prog(const ph_prog *p, int n_in, const int *in_dt, bool allow_index) {
  if (!p) return ph::set_error(PH_E_ARGUMENT, "null program");
  if (p->n_insn < 1 || p->n_insn > PH_MAX_INSN)
    return ph::set_error(PH_E_ARGUMENT, "program length %d", p->n_insn);
  if (p->n_in != n_in) return ph::set_error(PH_E_ARGUMENT, "program expects %d inputs, got %d", p->n_in, n_in);
  int rdt[PH_NREG];
  for (int i = 0; i < PH_NREG; ++i) rdt[i] = -1;
  for (int k = 0; k < p->n_insn; ++k) {
    const ph_insn &in = p->insn[k];
    if (in.dst >= PH_NREG) return ph::set_error(PH_E_ARGUMENT, "insn %d: register %d", k, in.dst);
    const int dt = in.dt;
    if (dt != PH_I32 && dt != PH_F32 && dt != PH_F64 && dt != PH_BOOL)
      return ph::set_error(PH_E_ARGUMENT, "insn %d: dtype %d", k, dt);
    auto reg = [&](int r) { return (r < PH_NREG) ? rdt[r] : -2; };
    int out_dt = dt;
    switch (in.op) {
      case PH_OP_LOAD:
        if (in.a >= n_in) return ph::set_error(PH_E_ARGUMENT, "insn %d: input %d", k, in.a);
        if (compute_dt(in_dt[in.a]) != dt)
          return ph::set_error(PH_E_ARGUMENT, "insn %d: input %d loads as %d, not %d", k, in.a,
                               compute_dt(in_dt[in.a]), dt);
        break;
      case PH_OP_CONST: break;
      case PH_OP_CAST:
        if (reg(in.a) != in.c) return ph::set_error(PH_E_ARGUMENT, "insn %d: cast source dtype", k);
        break;
      case PH_OP_ADD: case PH_OP_SUB: case PH_OP_MUL:
        if (dt == PH_BOOL) return ph::set_error(PH_E_ARGUMENT, "insn %d: arithmetic on Bool", k);
        if (reg(in.a) != dt || reg(in.b) != dt) return ph::set_error(PH_E_ARGUMENT, "insn %d: operand dtype", k);
        break;
      case PH_OP_WADD: case PH_OP_WSUB: case PH_OP_WMUL: case PH_OP_IPOW: case PH_OP_WIPOW:
      case PH_OP_BNOT:
        if (dt != PH_I32) return ph::set_error(PH_E_ARGUMENT, "insn %d: Int32-only op", k);
        if (reg(in.a) != dt || (in.op != PH_OP_BNOT && reg(in.b) != dt))
          return ph::set_error(PH_E_ARGUMENT, "insn %d: operand dtype", k);
        break;
      case PH_OP_DIV: case PH_OP_POW: case PH_OP_SQRT:
        if (!is_float(dt)) return ph::set_error(PH_E_ARGUMENT, "insn %d: float-only op", k);
        if (reg(in.a) != dt || (in.op != PH_OP_SQRT && reg(in.b) != dt))
          return ph::set_error(PH_E_ARGUMENT, "insn %d: operand dtype", k);
        break;
      case PH_OP_FLOORDIV: case PH_OP_MOD:
        if (dt == PH_BOOL) return ph::set_error(PH_E_ARGUMENT, "insn %d: // or %% on Bool", k);
        if (reg(in.a) != dt || reg(in.b) != dt) return ph::set_error(PH_E_ARGUMENT, "insn %d: operand dtype", k);
        break;
      case PH_OP_POWI:
        if (!is_float(dt) || reg(in.a) != dt || reg(in.b) != PH_I32)
          return ph::set_error(PH_E_ARGUMENT, "insn %d: powi operands", k);
        break;
      case PH_OP_NEG: case PH_OP_ABS:
        if (dt == PH_BOOL || reg(in.a) != dt) return ph::set_error(PH_E_ARGUMENT, "insn %d: operand dtype", k);
        break;
      case PH_OP_AND: case PH_OP_OR: case PH_OP_XOR:
        if (!(dt == PH_I32 || dt == PH_BOOL) || reg(in.a) != dt || reg(in.b) != dt)
          return ph::set_error(PH_E_ARGUMENT, "insn %d: bitwise operands", k);
        break;
      case PH_OP_GT: case PH_OP_LT: case PH_OP_GE: case PH_OP_LE: case PH_OP_EQ: case PH_OP_NE:
        if (reg(in.a) != dt || reg(in.b) != dt) return ph::set_error(PH_E_ARGUMENT, "insn %d: comparison operands", k);
        out_dt = PH_BOOL;
        break;
      case PH_OP_CMP:
        if (reg(in.a) != dt || reg(in.b) != dt) return ph::set_error(PH_E_ARGUMENT, "insn %d: <=> operands", k);
        out_dt = PH_I32;
        break;
      case PH_OP_SELECT:
        if (reg(in.c) != PH_BOOL || reg(in.a) != dt || reg(in.b) != dt)
          return ph::set_error(PH_E_ARGUMENT, "insn %d: select operands", k);
        break;
      case PH_OP_INDEX:
        if (!allow_index) return ph::set_error(PH_E_ARGUMENT, "insn %d: element index outside a map", k);
        if (dt != PH_I32) return ph::set_error(PH_E_ARGUMENT, "insn %d: element index is Int32", k);
        break;
      default:
        return ph::set_error(PH_E_ARGUMENT, "insn %d: opcode %d", k, in.op);
    }
    rdt[in.dst] = out_dt;
  }
  if (p->out_reg < 0 || p->out_reg >= PH_NREG || rdt[p->out_reg] < 0)
    return ph::set_error(PH_E_ARGUMENT, "program result register %d undefined", p->out_reg);
  if (rdt[p->out_reg] != p->out_dtype)
    return ph::set_error(PH_E_ARGUMENT, "program result dtype %d, declared %d", rdt[p->out_reg], p->out_dtype);
  return PH_OK;
}

bool prog_uses_index(const ph_prog *p) {
  for (int k = 0; k < p->n_insn; ++k)
    if (p->insn[k].op == PH_OP_INDEX) return true;
  return false;
}

// Whether the value the reduction's maxima take can never be a signalling
// NaN: it is widened from Float32 to Float64 on the way (a conversion), or
// the program's result comes out of a floating-point operation that the
// compiler cannot fold into a pass-through of its operand (no identity
// constant: x * 1, x / 1, x - 0, x + -0 are folded).  Loads, constants,
// selects and sign changes can pass a signalling NaN through unchanged.
bool max_operand_quiet(const ph_prog *p, bool widened) {
  if (widened) return true;
  int def[PH_NREG];
  for (int r = 0; r < PH_NREG; ++r) def[r] = -1;
  for (int k = 0; k < p->n_insn; ++k)
    if (p->insn[k].dst < PH_NREG) def[p->insn[k].dst] = k;
  auto const_val = [&](int reg, int before, double *v) {
    int d = -1;
    for (int k = 0; k < before; ++k)
      if (p->insn[k].dst == reg) d = k;
    if (d < 0 || p->insn[d].op != PH_OP_CONST) return false;
    const ph_insn &c = p->insn[d];
    *v = c.dt == PH_I32 ? (double)c.imm.i : c.imm.f;
    return true;
  };
  const int k = p->out_reg < PH_NREG ? def[p->out_reg] : -1;
  if (k < 0) return false;
  const ph_insn &in = p->insn[k];
  if (in.dt != PH_F32 && in.dt != PH_F64) return false;
  double va = 0, vb = 0;
  const bool ca = const_val(in.a, k, &va), cb = const_val(in.b, k, &vb);
  switch (in.op) {
    case PH_OP_ADD: return !(ca && va == 0.0 && std::signbit(va)) && !(cb && vb == 0.0 && std::signbit(vb));
    case PH_OP_SUB: return !(cb && vb == 0.0 && !std::signbit(vb)) && !(ca && va == 0.0 && std::signbit(va));
    case PH_OP_MUL: return !(ca && std::fabs(va) == 1.0) && !(cb && std::fabs(vb) == 1.0);
    case PH_OP_DIV: return !(cb && std::fabs(vb) == 1.0);
    case PH_OP_FLOORDIV:
    case PH_OP_MOD:
    case PH_OP_SQRT: return true;
    case PH_OP_CAST: return in.c != in.dt;
    default: return false;
  }
}

// Emits the statements computing the program for element `e`; inputs are
// the expressions `inexpr[j]` (already of their compute type).  Returns the
// name of the result variable.
// Exact multiply-add fusion.  A Float64 product x * 2^k, where x is a
// Float32 or Int32 value widened to Float64 and 2^k a constant, is exact
// (x has at most 31 significant bits and, for the k accepted below, the
// product neither overflows nor loses bits below 2^-1074).  An add or
// subtract of that product is then computed as one fused multiply-add with
// the same single rounding as the two operations, so the fusion changes no
// result.  Returns the exponent k of a constant 2^k in the accepted range for
// the source type, or 0 with *ok false.
static int exact_pow2(double c, int src_dt, bool *ok) {
  *ok = false;
  if (!(c > 0.0) || c != c) return 0;
  int e = 0;
  const double m = std::frexp(c, &e);
  if (m != 0.5) return 0;  // not a power of two
  const int k = e - 1;
  const bool fits = src_dt == PH_F32 ? (k >= -900 && k <= 890) : src_dt == PH_I32 ? (k >= -1000 && k <= 990) : false;
  *ok = fits;
  return k;
}

std::string emit_prog(const ph_prog *p, const std::vector<std::string> &inexpr, std::string *code,
                      const std::string &sfx) {
  std::string regv[PH_NREG];
  // per register: the Float64 value widened from (dtype, or -1), a Float64
  // constant, and an exact product x * 2^k (its x and the constant)
  int widened[PH_NREG];
  bool isconst[PH_NREG];
  double cval[PH_NREG];
  std::string px[PH_NREG], pk[PH_NREG];
  for (int r = 0; r < PH_NREG; ++r) {
    widened[r] = -1;
    isconst[r] = false;
    cval[r] = 0.0;
  }
  for (int k = 0; k < p->n_insn; ++k) {
    const ph_insn &in = p->insn[k];
    const int dt = in.dt;
    const std::string v = fmt("v%d%s", k, sfx.c_str());
    const std::string A = in.a < PH_NREG ? regv[in.a] : "";
    const std::string B = in.b < PH_NREG ? regv[in.b] : "";
    const std::string C = in.c < PH_NREG ? regv[in.c] : "";
    const char *T = ctype(dt);
    std::string s;
    // register facts after this instruction
    int nwid = -1;
    bool ncon = false;
    double nval = 0.0;
    std::string npx, npk;
    if (in.op == PH_OP_CAST && dt == PH_F64 && (in.c == PH_F32 || in.c == PH_I32)) nwid = in.c;
    if (in.op == PH_OP_CONST && dt == PH_F64) {
      ncon = true;
      nval = in.imm.f;
    }
    if (in.op == PH_OP_MUL && dt == PH_F64 && in.a < PH_NREG && in.b < PH_NREG) {
      bool ok = false;
      if (widened[in.a] >= 0 && isconst[in.b]) {
        exact_pow2(cval[in.b], widened[in.a], &ok);
        if (ok) { npx = A; npk = B; }
      } else if (widened[in.b] >= 0 && isconst[in.a]) {
        exact_pow2(cval[in.a], widened[in.b], &ok);
        if (ok) { npx = B; npk = A; }
      }
    }
    if ((in.op == PH_OP_ADD || in.op == PH_OP_SUB) && dt == PH_F64 && in.a < PH_NREG && in.b < PH_NREG &&
        (!px[in.a].empty() || !px[in.b].empty())) {
      std::string f;
      if (!px[in.a].empty())  // (x * 2^k) +/- B
        f = fmt("__builtin_fma(%s, %s, %s%s)", px[in.a].c_str(), pk[in.a].c_str(), in.op == PH_OP_SUB ? "-" : "",
                B.c_str());
      else if (in.op == PH_OP_ADD)  // A + (x * 2^k)
        f = fmt("__builtin_fma(%s, %s, %s)", px[in.b].c_str(), pk[in.b].c_str(), A.c_str());
      else  // A - (x * 2^k)
        f = fmt("__builtin_fma(-%s, %s, %s)", px[in.b].c_str(), pk[in.b].c_str(), A.c_str());
      *code += "        const double " + v + " = " + f + ";  // exact fusion (see exact_pow2)\n";
      regv[in.dst] = v;
      widened[in.dst] = -1;
      isconst[in.dst] = false;
      px[in.dst].clear();
      pk[in.dst].clear();
      continue;
    }
    switch (in.op) {
      case PH_OP_LOAD: s = fmt("const %s %s = %s;", T, v.c_str(), inexpr[in.a].c_str()); break;
      case PH_OP_CONST: s = fmt("const %s %s = %s;", T, v.c_str(), const_expr(dt, in).c_str()); break;
      case PH_OP_CAST: s = fmt("const %s %s = (%s)(%s);", T, v.c_str(), T, A.c_str()); break;
      case PH_OP_ADD: case PH_OP_SUB: case PH_OP_MUL: {
        const char *o = in.op == PH_OP_ADD ? "+" : in.op == PH_OP_SUB ? "-" : "*";
        if (dt == PH_I32)
          s = fmt("int %s; { const long long t_ = (long long)%s %s (long long)%s; fl |= (t_ != (long long)(int)t_) ? 1 : 0; %s = (int)t_; }",
                  v.c_str(), A.c_str(), o, B.c_str(), v.c_str());
        else
          s = fmt("const %s %s = %s %s %s;", T, v.c_str(), A.c_str(), o, B.c_str());
        break;
      }
      case PH_OP_WADD: case PH_OP_WSUB: case PH_OP_WMUL: {
        const char *o = in.op == PH_OP_WADD ? "+" : in.op == PH_OP_WSUB ? "-" : "*";
        s = fmt("const int %s = (int)((unsigned)%s %s (unsigned)%s);", v.c_str(), A.c_str(), o, B.c_str());
        break;
      }
      case PH_OP_DIV: s = fmt("const %s %s = %s / %s;", T, v.c_str(), A.c_str(), B.c_str()); break;
      case PH_OP_FLOORDIV:
        if (dt == PH_I32)
          s = fmt("const int %s = ph_ifloordiv(%s, %s, fl);", v.c_str(), A.c_str(), B.c_str());
        else
          s = fmt("const %s %s = ph_floor(%s / %s);", T, v.c_str(), A.c_str(), B.c_str());
        break;
      case PH_OP_MOD:
        if (dt == PH_I32)
          s = fmt("const int %s = ph_imod(%s, %s, fl);", v.c_str(), A.c_str(), B.c_str());
        else
          s = fmt("const %s %s = ph_fmod_floor(%s, %s);", T, v.c_str(), A.c_str(), B.c_str());
        break;
      case PH_OP_POWI: s = fmt("const %s %s = ph_powi(%s, %s);", T, v.c_str(), A.c_str(), B.c_str()); break;
      case PH_OP_IPOW: s = fmt("const int %s = ph_ipow(%s, %s, fl, true);", v.c_str(), A.c_str(), B.c_str()); break;
      case PH_OP_WIPOW: s = fmt("const int %s = ph_ipow(%s, %s, fl, false);", v.c_str(), A.c_str(), B.c_str()); break;
      case PH_OP_POW: s = fmt("const %s %s = ph_pow(%s, %s);", T, v.c_str(), A.c_str(), B.c_str()); break;
      case PH_OP_NEG:
        if (dt == PH_I32)
          s = fmt("const int %s = ph_ineg(%s, fl);", v.c_str(), A.c_str());
        else
          s = fmt("const %s %s = -%s;", T, v.c_str(), A.c_str());
        break;
      case PH_OP_BNOT: s = fmt("const int %s = ~%s;", v.c_str(), A.c_str()); break;
      case PH_OP_AND: case PH_OP_OR: case PH_OP_XOR: {
        const char *o = in.op == PH_OP_AND ? "&" : in.op == PH_OP_OR ? "|" : "^";
        if (dt == PH_BOOL)
          s = fmt("const bool %s = (bool)((int)%s %s (int)%s);", v.c_str(), A.c_str(), o, B.c_str());
        else
          s = fmt("const int %s = %s %s %s;", v.c_str(), A.c_str(), o, B.c_str());
        break;
      }
      case PH_OP_GT: case PH_OP_LT: case PH_OP_GE: case PH_OP_LE: case PH_OP_EQ: case PH_OP_NE: {
        static const char *ops[] = {">", "<", ">=", "<=", "==", "!="};
        s = fmt("const bool %s = %s %s %s;", v.c_str(), A.c_str(), ops[in.op - PH_OP_GT], B.c_str());
        break;
      }
      case PH_OP_CMP: s = fmt("const int %s = (int)(%s > %s) - (int)(%s < %s);", v.c_str(), A.c_str(), B.c_str(), A.c_str(), B.c_str()); break;
      case PH_OP_SQRT: s = fmt("const %s %s = ph_sqrt(%s);", T, v.c_str(), A.c_str()); break;
      case PH_OP_ABS:
        if (dt == PH_I32)
          s = fmt("const int %s = ph_iabs(%s, fl);", v.c_str(), A.c_str());
        else
          s = fmt("const %s %s = ph_fabs(%s);", T, v.c_str(), A.c_str());
        break;
      case PH_OP_SELECT: s = fmt("const %s %s = %s ? %s : %s;", T, v.c_str(), C.c_str(), A.c_str(), B.c_str()); break;
      case PH_OP_INDEX: s = fmt("const int %s = (int)ph_lidx;", v.c_str()); break;
    }
    *code += "        " + s + "\n";
    regv[in.dst] = v;
    widened[in.dst] = nwid;
    isconst[in.dst] = ncon;
    cval[in.dst] = nval;
    px[in.dst] = npx;
    pk[in.dst] = npk;
  }
  return regv[p->out_reg];
}

extern const char *const kPrelude;
const char *const kPrelude = R"PH(
typedef long long i64;
__device__ __forceinline__ float ph_ld(const unsigned short *p, i64 i) { return __uint_as_float(((unsigned)p[i]) << 16); }
__device__ __forceinline__ bool ph_ld(const unsigned char *p, i64 i) { return p[i] != 0; }
__device__ __forceinline__ int ph_ld(const int *p, i64 i) { return p[i]; }
__device__ __forceinline__ float ph_ld(const float *p, i64 i) { return p[i]; }
__device__ __forceinline__ double ph_ld(const double *p, i64 i) { return p[i]; }
__device__ __forceinline__ void ph_st(unsigned char *p, i64 i, bool v) { p[i] = v ? 1 : 0; }
// results are written once and not read again by the kernel: non-temporal
__device__ __forceinline__ void ph_st(int *p, i64 i, int v) { __builtin_nontemporal_store(v, p + i); }
__device__ __forceinline__ void ph_st(float *p, i64 i, float v) { __builtin_nontemporal_store(v, p + i); }
__device__ __forceinline__ void ph_st(double *p, i64 i, double v) { __builtin_nontemporal_store(v, p + i); }
__device__ __forceinline__ float ph_floor(float x) { return __builtin_floorf(x); }
__device__ __forceinline__ double ph_floor(double x) { return __builtin_floor(x); }
__device__ __forceinline__ float ph_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double ph_sqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float ph_fabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double ph_fabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float ph_pow(float a, float b) { return __builtin_powf(a, b); }
__device__ __forceinline__ double ph_pow(double a, double b) { return __builtin_pow(a, b); }
__device__ __forceinline__ float ph_fmod_floor(float a, float b) { const float q = __builtin_floorf(a / b); const float m = b * q; return a - m; }
__device__ __forceinline__ double ph_fmod_floor(double a, double b) { const double q = __builtin_floor(a / b); const double m = b * q; return a - m; }
template <typename F> __device__ __forceinline__ F ph_powi(F a, int b) {
  const bool recip = b < 0; long long e = b; if (e < 0) e = -e; F r = (F)1;
  for (;;) { if (e & 1) r *= a; e /= 2; if (e == 0) break; a *= a; }
  return recip ? (F)1 / r : r;
}
__device__ __forceinline__ int ph_ipow(int base, int e, int &fl, bool checked) {
  if (e < 0) { fl |= 4; return 0; }
  long long result = 1, k = base; long long ex = e;
  while (ex > 0) {
    if (ex & 1) { const long long t = result * k; if (checked && t != (long long)(int)t) fl |= 1; result = (int)t; }
    ex >>= 1;
    if (ex > 0) { const long long t = k * k; if (checked && t != (long long)(int)t) fl |= 1; k = (int)t; }
  }
  return (int)result;
}
__device__ __forceinline__ int ph_ifloordiv(int a, int b, int &fl) {
  if (b == 0) { fl |= 2; return 0; }
  if (a == (-2147483647 - 1) && b == -1) { fl |= 1; return a; }
  int q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1; return q;
}
__device__ __forceinline__ int ph_imod(int a, int b, int &fl) {
  if (b == 0) { fl |= 2; return 0; }
  if (b == -1) return 0;
  int m = a % b; if (m != 0 && ((m < 0) != (b < 0))) m += b; return m;
}
__device__ __forceinline__ int ph_ineg(int a, int &fl) { if (a == (-2147483647 - 1)) fl |= 1; return (int)(0u - (unsigned)a); }
__device__ __forceinline__ int ph_iabs(int a, int &fl) { if (a == (-2147483647 - 1)) fl |= 1; return a < 0 ? (int)(0u - (unsigned)a) : a; }
)PH";

// Whether any element can raise a flag.  Int32 // and % by a constant other
// than 0 and -1 cannot (the divisor is known when the program is built, so
// neither division by zero nor the MIN / -1 overflow can occur): a
// coordinate computed from the element index then costs no read-back.
bool prog_may_flag(const ph_prog *p) {
  bool isc[PH_NREG] = {false};
  int64_t cv[PH_NREG] = {0};
  for (int k = 0; k < p->n_insn; ++k) {
    const ph_insn &in = p->insn[k];
    if (in.op == PH_OP_IPOW || in.op == PH_OP_WIPOW) return true;
    if (in.dt == PH_I32 && (in.op == PH_OP_FLOORDIV || in.op == PH_OP_MOD)) {
      const bool safe = in.b < PH_NREG && isc[in.b] && cv[in.b] != 0 && cv[in.b] != -1;
      if (!safe) return true;
    }
    if (in.dt == PH_I32 && (in.op == PH_OP_ADD || in.op == PH_OP_SUB || in.op == PH_OP_MUL ||
                            in.op == PH_OP_NEG || in.op == PH_OP_ABS))
      return true;
    if (in.dst < PH_NREG) {
      isc[in.dst] = in.op == PH_OP_CONST && in.dt == PH_I32;
      cv[in.dst] = in.imm.i;
    }
  }
  return false;
}

}  // namespace gen
}  // namespace ph

namespace {

using namespace ph::gen;

// ---------------------------------------------------------------------------
// Map
// ---------------------------------------------------------------------------
struct MapArgs {
  const void *in[PH_MAX_INPUTS];
  int64_t in_off[PH_MAX_INPUTS];
  void *out;
  int64_t out_off;
  int *flags;
};

constexpr int kBlock = 256;

// Elements per lane of the map kernel: the most (up to 8) that keep a
// lane's loads within 48 bytes.  Measured on 16384^2 Float64
// (a + b.tile) * c: 2 per lane (48 bytes) 1.02 ms, 4 1.10 ms, 8 1.17 ms; on
// 8192^2 Int32 maps: x.eq(y) 0.100 ms at 4 against 0.106 at 8, z[mask] = y
// 0.143 against 0.155, and z[mask] = 7 (5 bytes) 0.098 ms at 8 against 0.124
// at 4.
int map_ept(int n_in, const int *in_dt) {
  int in_bytes = 0;
  for (int j = 0; j < n_in; ++j) in_bytes += (int)ph::dtype_size(in_dt[j]);
  int ept = 8;
  while (ept > 1 && ept * in_bytes > 48) ept /= 2;
  return ept;
}

std::string map_source(const ph_prog *p, int n_in, const int *in_dt, const ph::Layout &L,
                       int out_dt, int ept, bool flat, int64_t *nblocks) {
  // operands 0..n_in-1 inputs, n_in output
  const int outer = L.rank > 0 ? L.rank - 1 : 0;
  const int64_t cols = L.rank > 0 ? L.shape[L.rank - 1] : 1;
  int64_t rows = 1;
  for (int k = 0; k < outer; ++k) rows *= L.shape[k];
  const int64_t tile = (int64_t)kBlock * ept;
  std::string s = kPrelude;
  s += fmt("struct MapArgs { const void *in[%d]; i64 in_off[%d]; void *out; i64 out_off; int *flags; };\n",
           PH_MAX_INPUTS, PH_MAX_INPUTS);
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_map_kernel(MapArgs A) {\n";
  s += "  int fl = 0;\n";
  auto inner_stride = [&](int op) { return L.rank > 0 ? L.stride[op][L.rank - 1] : 0; };
  auto decode = [&](const char *rowvar) {
    std::string d;
    d += fmt("    i64 rr = %s;\n", rowvar);
    for (int o = 0; o <= n_in; ++o)
      d += fmt("    i64 base%d = %s;\n", o, o < n_in ? fmt("A.in_off[%d]", o).c_str() : "A.out_off");
    for (int k = outer - 1; k >= 0; --k) {
      d += fmt("    { const i64 q = rr / %s; const i64 c = rr - q * %s; rr = q;\n", lit_i64(L.shape[k]).c_str(),
               lit_i64(L.shape[k]).c_str());
      for (int o = 0; o <= n_in; ++o)
        if (L.stride[o][k] != 0) d += fmt("      base%d += c * %s;\n", o, lit_i64(L.stride[o][k]).c_str());
      d += "    }\n";
    }
    return d;
  };
  std::vector<std::string> inexpr;
  for (int j = 0; j < n_in; ++j) inexpr.push_back(fmt("x%d[e]", j));
  std::string body;
  std::string res = emit_prog(p, inexpr, &body, "");
  // the element's lexicographic index: collapse() keeps the logical order
  // (it drops extent-1 axes and merges neighbours), so it is row * cols + c
  const bool uses_index = prog_uses_index(p);
  if (!flat) {
    const int64_t ntiles = (cols + tile - 1) / tile;
    *nblocks = rows * ntiles;
    s += fmt("  for (i64 b = blockIdx.x; b < %s; b += gridDim.x) {\n", lit_i64(*nblocks).c_str());
    s += fmt("    const i64 row = b / %s; const i64 tile = b - row * %s;\n", lit_i64(ntiles).c_str(),
             lit_i64(ntiles).c_str());
    s += decode("row");
    s += fmt("    const i64 c0 = tile * %s + threadIdx.x;\n", lit_i64(tile).c_str());
    for (int j = 0; j < n_in; ++j) s += fmt("    %s x%d[%d];\n", ctype(compute_dt(in_dt[j])), j, ept);
    s += fmt("#pragma unroll\n    for (int e = 0; e < %d; ++e) {\n", ept);
    s += fmt("      const i64 c = c0 + (i64)e * 256;\n      if (c < %s) {\n", lit_i64(cols).c_str());
    for (int j = 0; j < n_in; ++j)
      s += fmt("        x%d[e] = ph_ld((const %s *)A.in[%d], base%d + c * %s);\n", j, stype(in_dt[j]), j, j,
               lit_i64(inner_stride(j)).c_str());
    s += "      }\n    }\n";
    s += fmt("#pragma unroll\n    for (int e = 0; e < %d; ++e) {\n", ept);
    s += fmt("      const i64 c = c0 + (i64)e * 256;\n      if (c < %s) {\n", lit_i64(cols).c_str());
    if (uses_index) s += fmt("        const i64 ph_lidx = row * %s + c;\n", lit_i64(cols).c_str());
    s += body;
    s += fmt("        ph_st((%s *)A.out, base%d + c * %s, %s);\n", stype(out_dt), n_in,
             lit_i64(inner_stride(n_in)).c_str(), res.c_str());
    s += "      }\n    }\n  }\n";
  } else {
    const int64_t total = rows * cols;
    *nblocks = (total + tile - 1) / tile;
    s += fmt("  for (i64 b = blockIdx.x; b < %s; b += gridDim.x) {\n", lit_i64(*nblocks).c_str());
    s += fmt("#pragma unroll\n    for (int e = 0; e < %d; ++e) {\n", ept);
    s += fmt("      const i64 g = b * %s + (i64)e * 256 + threadIdx.x;\n", lit_i64(tile).c_str());
    s += fmt("      if (g < %s) {\n", lit_i64(total).c_str());
    s += fmt("        const i64 row = g / %s; const i64 c = g - row * %s;\n", lit_i64(cols).c_str(),
             lit_i64(cols).c_str());
    std::string d = decode("row");
    s += d;
    for (int j = 0; j < n_in; ++j)
      s += fmt("        %s x%d[1]; x%d[0] = ph_ld((const %s *)A.in[%d], base%d + c * %s);\n",
               ctype(compute_dt(in_dt[j])), j, j, stype(in_dt[j]), j, j, lit_i64(inner_stride(j)).c_str());
    s += "        const int e0 = 0; (void)e0;\n";
    std::string body0;
    std::vector<std::string> inexpr0;
    for (int j = 0; j < n_in; ++j) inexpr0.push_back(fmt("x%d[0]", j));
    std::string res0 = emit_prog(p, inexpr0, &body0, "");
    if (uses_index) s += "        const i64 ph_lidx = g;\n";
    s += body0;
    s += fmt("        ph_st((%s *)A.out, base%d + c * %s, %s);\n", stype(out_dt), n_in,
             lit_i64(inner_stride(n_in)).c_str(), res0.c_str());
    s += "      }\n    }\n  }\n";
  }
  s += "  if (fl) atomicOr(A.flags, fl);\n}\n";
  return s;
}

// Map whose transposed inputs (bit j of `tmask`) are contiguous along axis
// `ax` of the collapsed layout while the output is contiguous along the
// last axis.  A workgroup owns a 64 x 64 tile of (ax, last) for one
// coordinate of the other axes: it reads the tile of each transposed input
// with lanes along `ax` into LDS ([last][ax], rows padded to 65 values),
// then computes with lanes along the last axis, reading the other inputs
// directly and the transposed ones from LDS, and writes the output.  So
// every operand moves whole cache lines and the transposed inputs need no
// pass of their own.
std::string map_tiled_source(const ph_prog *p, int n_in, const int *in_dt, const ph::Layout &L, int out_dt, int ax,
                             unsigned tmask, int64_t *nblocks) {
  const int bx = L.rank - 1;
  const int64_t na = L.shape[ax], nb = L.shape[bx];
  const int64_t ta = (na + 63) / 64, tb = (nb + 63) / 64;
  int64_t others = 1;
  for (int k = 0; k < L.rank; ++k)
    if (k != ax && k != bx) others *= L.shape[k];
  *nblocks = others * ta * tb;
  // the tile row a wave takes at step r of 16: every fourth row (sixteen
  // consecutive rows per wave measured 0.20 ms against 0.153 for an Int32
  // 8192^2 map)
  auto ROW = [&](const std::string &r) { return "(wave + 4 * " + r + ")"; };
  std::string s = kPrelude;
  s += fmt("struct MapArgs { const void *in[%d]; i64 in_off[%d]; void *out; i64 out_off; int *flags; };\n",
           PH_MAX_INPUTS, PH_MAX_INPUTS);
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_map_tiled(MapArgs A) {\n";
  s += "  int fl = 0;\n  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n";
  for (int j = 0; j < n_in; ++j)
    if (tmask >> j & 1u) s += fmt("  __shared__ %s lds%d[64][65];\n", ctype(compute_dt(in_dt[j])), j);
  s += fmt("  for (i64 b = blockIdx.x; b < %s; b += gridDim.x) {\n", lit_i64(*nblocks).c_str());
  s += fmt("    i64 o = b / %s; const i64 rem = b - o * %s;\n", lit_i64(ta * tb).c_str(), lit_i64(ta * tb).c_str());
  s += fmt("    const i64 ib = rem / %s, ia = rem - ib * %s;\n", lit_i64(ta).c_str(), lit_i64(ta).c_str());
  for (int o = 0; o <= n_in; ++o)
    s += fmt("    i64 base%d = %s;\n", o, o < n_in ? fmt("A.in_off[%d]", o).c_str() : "A.out_off");
  for (int k = L.rank - 1; k >= 0; --k) {
    if (k == ax || k == bx) continue;
    s += fmt("    { const i64 q = o / %s; const i64 c = o - q * %s; o = q;\n", lit_i64(L.shape[k]).c_str(),
             lit_i64(L.shape[k]).c_str());
    for (int o = 0; o <= n_in; ++o)
      if (L.stride[o][k] != 0) s += fmt("      base%d += c * %s;\n", o, lit_i64(L.stride[o][k]).c_str());
    s += "    }\n";
  }
  s += "    const i64 a0 = ia * 64, b0 = ib * 64;\n";
  // phase 1: the transposed inputs into LDS, lanes along ax
  s += "    {\n      const i64 ar = a0 + lane;\n#pragma unroll\n      for (int r = 0; r < 16; ++r) {\n";
  s += fmt((std::string("        const i64 bb = b0 + " + ROW("r") + ";\n        if (ar < %s && bb < %s) {\n")).c_str(), lit_i64(na).c_str(),
           lit_i64(nb).c_str());
  for (int j = 0; j < n_in; ++j)
    if (tmask >> j & 1u)
      s += fmt((std::string("          lds%d[" + ROW("r") + "][lane] = ph_ld((const %s *)A.in[%d], base%d + ar * %s + bb * %s);\n")).c_str(), j,
               stype(in_dt[j]), j, j, lit_i64(L.stride[j][ax]).c_str(), lit_i64(L.stride[j][bx]).c_str());
  s += "        }\n      }\n    }\n    __syncthreads();\n";
  // phase 2: compute with lanes along the last axis.  A whole tile runs
  // without bounds checks, `batch` rows at a time: the loads of the other
  // inputs for those rows, then their results.  One row at a time measured
  // best (0.154 ms for an Int32 8192^2 map against 0.160 with a check per
  // element, and 0.18-0.23 ms for 2-16 rows: a tile's rows lie a row pitch
  // apart, 32 KiB there, and more of them in flight at once only queue on
  // the same memory channels).  A tile at the edge checks each element.
  std::vector<std::string> inexpr;
  for (int j = 0; j < n_in; ++j) inexpr.push_back(fmt("x%d", j));
  std::string body;
  const std::string res = emit_prog(p, inexpr, &body, "");
  int batch = 1;
  if (const char *e = std::getenv("PH_TMAP_BATCH")) batch = std::atoi(e);  // tuning override
  if (batch != 1 && batch != 2 && batch != 4 && batch != 8 && batch != 16) batch = 1;
  s += fmt("    if (a0 + 64 <= %s && b0 + 64 <= %s) {\n      const i64 bc = b0 + lane;\n", lit_i64(na).c_str(),
           lit_i64(nb).c_str());
  s += fmt("#pragma unroll\n      for (int r0 = 0; r0 < 16; r0 += %d) {\n", batch);
  for (int j = 0; j < n_in; ++j) {
    if (tmask >> j & 1u) continue;
    const std::string f = std::string("        %s y%d[%d];\n#pragma unroll\n        for (int u = 0; u < %d; ++u) "
                                      "y%d[u] = ph_ld((const %s *)A.in[%d], base%d + (a0 + ") +
                          ROW("(r0 + u)") + ") * %s + bc * %s);\n";
    s += fmt(f.c_str(),
             ctype(compute_dt(in_dt[j])), j, batch, batch, j, stype(in_dt[j]), j, j, lit_i64(L.stride[j][ax]).c_str(),
             lit_i64(L.stride[j][bx]).c_str());
  }
  s += fmt("#pragma unroll\n        for (int u = 0; u < %d; ++u) {\n          const int r = r0 + u;\n", batch);
  s += "          const i64 aa = a0 + " + ROW("r") + ";\n";
  for (int j = 0; j < n_in; ++j) {
    if (tmask >> j & 1u)
      s += fmt((std::string("          const %s x%d = lds%d[lane][" + ROW("r") + "];\n")).c_str(), ctype(compute_dt(in_dt[j])), j, j);
    else
      s += fmt("          const %s x%d = y%d[u];\n", ctype(compute_dt(in_dt[j])), j, j);
  }
  s += body;
  s += fmt("          ph_st((%s *)A.out, base%d + aa * %s + bc * %s, %s);\n        }\n      }\n    } else\n", stype(out_dt),
           n_in, lit_i64(L.stride[n_in][ax]).c_str(), lit_i64(L.stride[n_in][bx]).c_str(), res.c_str());
  s += "    {\n      const i64 bc = b0 + lane;\n#pragma unroll 4\n      for (int r = 0; r < 16; ++r) {\n";
  s += fmt((std::string("        const i64 aa = a0 + " + ROW("r") + ";\n        if (aa < %s && bc < %s) {\n")).c_str(), lit_i64(na).c_str(),
           lit_i64(nb).c_str());
  for (int j = 0; j < n_in; ++j) {
    if (tmask >> j & 1u)
      s += fmt((std::string("          const %s x%d = lds%d[lane][" + ROW("r") + "];\n")).c_str(), ctype(compute_dt(in_dt[j])), j, j);
    else
      s += fmt("          const %s x%d = ph_ld((const %s *)A.in[%d], base%d + aa * %s + bc * %s);\n",
               ctype(compute_dt(in_dt[j])), j, stype(in_dt[j]), j, j, lit_i64(L.stride[j][ax]).c_str(),
               lit_i64(L.stride[j][bx]).c_str());
  }
  s += body;
  s += fmt("          ph_st((%s *)A.out, base%d + aa * %s + bc * %s, %s);\n", stype(out_dt), n_in,
           lit_i64(L.stride[n_in][ax]).c_str(), lit_i64(L.stride[n_in][bx]).c_str(), res.c_str());
  s += "        }\n      }\n    }\n    __syncthreads();  // the next tile reuses the LDS\n  }\n";
  s += "  if (fl) atomicOr(A.flags, fl);\n}\n";
  return s;
}

// The axis along which every input of `tmask` is contiguous in the collapsed
// layout, if the tiled map applies: it is not the last axis, the output is
// contiguous along the last axis, and both extents are 16 or more.
int tiled_map_axis(const ph::Layout &L, int n_in, unsigned tmask) {
  if (L.rank < 2) return -1;
  const int bx = L.rank - 1;
  const int64_t so = L.stride[n_in][bx];
  if (!(so == 1 || so == -1) || L.shape[bx] < 16) return -1;
  int ax = -1;
  for (int j = 0; j < n_in; ++j) {
    if (!(tmask >> j & 1u)) continue;
    int aj = -1;
    for (int k = 0; k < bx; ++k)
      if (L.stride[j][k] == 1 || L.stride[j][k] == -1) {
        aj = k;
        break;
      }
    if (aj < 0 || (ax >= 0 && aj != ax)) return -1;
    ax = aj;
  }
  return (ax >= 0 && L.shape[ax] >= 16) ? ax : -1;
}

int64_t grid_for(int64_t blocks) {
  const int64_t cap = 1 << 20;
  return blocks < 1 ? 1 : (blocks > cap ? cap : blocks);
}

// The flag word is cleared by a one-thread kernel (not a memset node, so the
// sequence stays valid inside a captured HIP graph) and read back only when
// the program contains an operation that can raise a flag.
int get_flags_buffer(int **dflags, bool clear) {
  void *p;
  int rc = ph::scratch(64, &p, 3);
  if (rc) return rc;
  *dflags = (int *)p;
  if (clear) return ph::clear_i32(*dflags, 1);
  return PH_OK;
}

int read_flags(int *dflags, bool may_flag, int32_t *flags_out) {
  if (!flags_out) return PH_OK;
  if (!may_flag) {
    *flags_out = 0;
    return PH_OK;
  }
  int v = 0;
  PH_HIP(hipMemcpyAsync(&v, dflags, sizeof(int), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  *flags_out = v;
  return PH_OK;
}

// Compiled fused kernels, keyed by everything their source depends on (the
// program's instructions, operand dtypes and baked layout), so repeated calls
// skip source generation.
struct FusedKernels {
  hipFunction_t k[3];
  int64_t nblocks;
};
std::map<std::string, FusedKernels> &fused_cache() {
  static std::map<std::string, FusedKernels> c;
  return c;
}

void key_add(std::string *k, const void *p, size_t n) { k->append((const char *)p, n); }

std::string prog_key(const ph_prog *p) {
  std::string k;
  key_add(&k, &p->n_insn, sizeof(int32_t) * 4);
  for (int i = 0; i < p->n_insn; ++i) {
    const ph_insn &in = p->insn[i];
    key_add(&k, &in.op, 6);
    key_add(&k, &in.imm, 8);
  }
  return k;
}

// Materialises staged (reshape / tile) inputs into contiguous scratch so the
// fused kernels only see affine operands.
// A transposed affine view: its innermost axis (of extent 16 or more) steps
// across cache lines while another axis of 16 or more elements is
// contiguous.  A kernel whose lanes run along the innermost axis would touch
// one line per element; one pass of the LDS-tiled transposing copy into
// contiguous scratch (the path ph_strided_copy takes for this layout) costs
// less than that.
bool transposed_layout(const ph_view_desc &d) {
  if (d.nstage != 0 || d.view.rank < 2) return false;
  int q = -1;
  for (int i = d.view.rank - 1; i >= 0; --i)
    if (d.view.shape[i] != 1) {
      q = i;
      break;
    }
  if (q < 0 || d.view.shape[q] < 16) return false;
  const int64_t sq = d.view.stride[q];
  if (sq == 0 || sq == 1 || sq == -1) return false;
  for (int i = 0; i < q; ++i)
    if (d.view.shape[i] >= 16 && (d.view.stride[i] == 1 || d.view.stride[i] == -1)) return true;
  return false;
}

// Whether the innermost axis (of extent above 1) of a view is contiguous.
bool inner_contiguous(const ph_view_desc &d) {
  for (int i = d.view.rank - 1; i >= 0; --i)
    if (d.view.shape[i] != 1) return d.view.stride[i] == 1 || d.view.stride[i] == -1;
  return true;
}

// A 2-D reduction whose every input is a transposed view (a unit stride
// along axis 0, a wider one along axis 1, both extents of 16 or more) and
// whose values are floating point can run in the storage's orientation:
// the view's row results are the storage's column results and the other
// way round.  Only the order of the additions changes, which the float
// sums' stated accuracy does not depend on.  Int32 sums keep the view's
// lexicographic order, which their running-total overflow check follows.
bool swap_transposed_2d(const ph_prog *prog, int n_in, const ph_view_desc *descs) {
  if (prog->out_dtype != PH_F32 && prog->out_dtype != PH_F64) return false;
  for (int j = 0; j < n_in; ++j) {
    const ph_view_desc &d = descs[j];
    if (d.nstage != 0 || d.view.rank != 2 || d.view.shape[0] < 16 || d.view.shape[1] < 16) return false;
    const int64_t r = d.view.stride[0], c = d.view.stride[1];
    if (!(r == 1 || r == -1) || c == 0 || c == 1 || c == -1) return false;
  }
  return true;
}

// The same view with its two axes exchanged.
ph_view_desc swapped_2d(const ph_view_desc &d) {
  ph_view_desc t = d;
  std::swap(t.view.shape[0], t.view.shape[1]);
  std::swap(t.view.stride[0], t.view.stride[1]);
  return t;
}

int materialize(const ph_view_desc *d, const void *src, int slot, ph_view_desc *nd, const void **np) {
  const int64_t n = ph::desc_count(d);
  void *buf;
  int rc = ph::scratch((size_t)(n > 0 ? n : 1) * ph::dtype_size(d->dtype), &buf, slot);
  if (rc) return rc;
  std::memset(nd, 0, sizeof(*nd));
  nd->dtype = d->dtype;
  nd->extent = n > 0 ? n : 1;
  nd->view.rank = d->view.rank;
  int64_t st = 1;
  for (int i = d->view.rank - 1; i >= 0; --i) {
    nd->view.shape[i] = d->view.shape[i];
    nd->view.stride[i] = st;
    st *= d->view.shape[i];
  }
  rc = ph_strided_copy(buf, nd, src, d, 0);
  if (rc) return rc;
  *np = buf;
  return PH_OK;
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
const char *kRedPrelude = R"PH(
struct SF { float s, c; double m; };  // the max is kept exactly, whatever the sum's precision
struct SD { double s, c, m; };
struct SI { i64 s, hi, lo; int m; int n; };
__device__ __forceinline__ void st_init(SF &a) { a.s = 0.f; a.c = 0.f; a.m = -__builtin_inf(); }
__device__ __forceinline__ void st_init(SD &a) { a.s = 0.0; a.c = 0.0; a.m = -__builtin_inf(); }
__device__ __forceinline__ void st_init(SI &a) { a.s = 0; a.hi = -(1LL << 62); a.lo = (1LL << 62); a.m = 0; a.n = 0; }
// A state from an FP64 partial sum and maximum: SF keeps the sum as an
// unevaluated pair of floats (s + c, about 48 significant bits).
__device__ __forceinline__ void st_from(SF &a, double s, double m) { a.s = (float)s; a.c = (float)(s - (double)a.s); a.m = m; }
__device__ __forceinline__ void st_from(SD &a, double s, double m) { a.s = s; a.c = 0.0; a.m = m; }
// V consecutive elements of one row (address aligned to min(V * size, 16)
// bytes), loaded with 16-byte (or, for fewer bytes, one) vector loads.
template <int V> __device__ __forceinline__ void ph_ldv(const float *p, float (&x)[V]) {
#pragma unroll
  for (int q = 0; q < V / 4; ++q) { const float4 t = ((const float4 *)p)[q]; x[4 * q] = t.x; x[4 * q + 1] = t.y; x[4 * q + 2] = t.z; x[4 * q + 3] = t.w; }
}
template <int V> __device__ __forceinline__ void ph_ldv(const double *p, double (&x)[V]) {
#pragma unroll
  for (int q = 0; q < V / 2; ++q) { const double2 t = ((const double2 *)p)[q]; x[2 * q] = t.x; x[2 * q + 1] = t.y; }
}
template <int V> __device__ __forceinline__ void ph_ldv(const int *p, int (&x)[V]) {
#pragma unroll
  for (int q = 0; q < V / 4; ++q) { const int4 t = ((const int4 *)p)[q]; x[4 * q] = t.x; x[4 * q + 1] = t.y; x[4 * q + 2] = t.z; x[4 * q + 3] = t.w; }
}
// BFloat16 and Bool elements stay packed in 32-bit words (2 and 4 per word,
// the lower address in the low bits) until they are used.
template <int V> __device__ __forceinline__ void ph_ldv(const unsigned short *p, unsigned (&w)[V / 2]) {
  if constexpr (V == 8) { const uint4 t = *(const uint4 *)p; w[0] = t.x; w[1] = t.y; w[2] = t.z; w[3] = t.w; }
  else { const uint2 t = *(const uint2 *)p; w[0] = t.x; w[1] = t.y; }
}
template <int V> __device__ __forceinline__ void ph_ldv(const unsigned char *p, unsigned (&w)[V / 4]) {
  if constexpr (V == 8) { const uint2 t = *(const uint2 *)p; w[0] = t.x; w[1] = t.y; }
  else { w[0] = *(const unsigned *)p; }
}
__device__ __forceinline__ float ph_bf(unsigned w, int h) { return __uint_as_float(h ? (w & 0xffff0000u) : (w << 16)); }
__device__ __forceinline__ bool ph_bb(unsigned w, int b) { return ((w >> (8 * b)) & 0xffu) != 0; }
__device__ __forceinline__ float ph_fmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double ph_fmax(double a, double b) { return __builtin_fmax(a, b); }
// a signalling NaN made quiet (any other value unchanged)
__device__ __forceinline__ float ph_quiet(float x) { return __builtin_canonicalizef(x); }
__device__ __forceinline__ double ph_quiet(double x) { return __builtin_canonicalize(x); }
// a = max(a, b) by the instruction itself, for operands that are never
// signalling NaNs (a quiet NaN operand yields the other one, as fmax)
__device__ __forceinline__ void ph_amax(double &a, double b) { __asm__ volatile("v_max_f64 %0, %0, %1" : "+v"(a) : "v"(b)); }
__device__ __forceinline__ void ph_amax(float &a, float b) { __asm__ volatile("v_max_f32 %0, %0, %1" : "+v"(a) : "v"(b)); }
// max of two values that are never signalling NaNs, by the instruction itself
__device__ __forceinline__ double ph_amax2(double a, double b) { double r; __asm__("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b)); return r; }
__device__ __forceinline__ float ph_amax2(float a, float b) { float r; __asm__("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b)); return r; }
template <typename F, bool MAX, bool RAW = false> __device__ __forceinline__ F ph_comb2(F a, F b) {
  if (MAX) return RAW ? ph_amax2(a, b) : ph_fmax(a, b);
  return a + b;
}
// Cross-lane moves.  v_permlane32_swap(a, b) exchanges lanes 32-63 of a with
// lanes 0-31 of b, v_permlane16_swap(a, b) the odd 16-lane rows of a with the
// even rows of b: with a = the lane's value for one row and b = its value for
// another, each lane ends up holding its own value and its partner's for the
// row it keeps, with no selects.  DPP moves serve the exchanges inside a row.
__device__ __forceinline__ void ph_swap32(float &a, float &b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  a = __uint_as_float(r[0]); b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void ph_swap16(float &a, float &b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  a = __uint_as_float(r[0]); b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void ph_swap32(double &a, double &b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a), ub = (unsigned long long)__double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ void ph_swap16(double &a, double &b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a), ub = (unsigned long long)__double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
// DPP controls: row_ror:8 (lane l reads l ^ 8 in its 16-lane row),
// row_half_mirror (l reads 7 - l in its 8-lane half row), quad_perm [2,3,0,1]
// (l ^ 2) and [1,0,3,2] (l ^ 1).
template <int C> __device__ __forceinline__ float ph_dpp(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), C, 0xF, 0xF, true));
}
template <int C> __device__ __forceinline__ double ph_dpp(double x) {
  const long long u = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, C, 0xF, 0xF, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)((unsigned long long)u >> 32), C, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Transposing butterfly over a wavefront: each lane holds 8 partial values
// (one per row of an 8-row group); at each exchange a lane keeps half of its
// rows and receives its partner's values for them, so after three exchanges
// (lane bits 5, 4, 3) lane l holds 8 lanes' totals for row (l >> 3) & 7, and
// a fixed three-level tree inside each 8-lane group (half mirror, then lane
// bits 1 and 0) completes the row over all 64 lanes.
// (RAW: maxima by the instruction itself, for values that are never
// signalling NaNs; values moved across lanes would otherwise be re-quieted)
template <typename F, bool MAX, bool RAW = false> __device__ __forceinline__ F tfold8(F (&v)[8]) {
  const int lane = threadIdx.x & 63;
  F a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3], b0 = v[4], b1 = v[5], b2 = v[6], b3 = v[7];
  ph_swap32(a0, b0); ph_swap32(a1, b1); ph_swap32(a2, b2); ph_swap32(a3, b3);
  // lanes 0-31: rows 0-3, lanes 32-63: rows 4-7
  F w0 = ph_comb2<F, MAX, RAW>(a0, b0), w1 = ph_comb2<F, MAX, RAW>(a1, b1);
  F w2 = ph_comb2<F, MAX, RAW>(a2, b2), w3 = ph_comb2<F, MAX, RAW>(a3, b3);
  ph_swap16(w0, w2); ph_swap16(w1, w3);
  const F y0 = ph_comb2<F, MAX, RAW>(w0, w2), y1 = ph_comb2<F, MAX, RAW>(w1, w3);
  const bool b8 = (lane & 8) != 0;
  const F keep = b8 ? y1 : y0, send = b8 ? y0 : y1;
  F x = ph_comb2<F, MAX, RAW>(keep, ph_dpp<0x128>(send));
  x = ph_comb2<F, MAX, RAW>(x, ph_dpp<0x141>(x));
  x = ph_comb2<F, MAX, RAW>(x, ph_dpp<0x4E>(x));
  x = ph_comb2<F, MAX, RAW>(x, ph_dpp<0xB1>(x));
  return x;
}
__device__ __forceinline__ void st_add(SI &a, i64 v) {
  a.s += v; a.hi = a.s > a.hi ? a.s : a.hi; a.lo = a.s < a.lo ? a.s : a.lo;
  a.m = (a.n == 0 || (int)v > a.m) ? (int)v : a.m; a.n = 1;
}
template <typename S> __device__ __forceinline__ S st_combf(const S &a, const S &b) {
  S r; const auto t = a.s + b.s; const auto bb = t - a.s; const auto err = (a.s - (t - bb)) + (b.s - bb);
  r.s = t; r.c = (a.c + b.c) + err; r.m = (b.m > a.m) ? b.m : a.m; return r;
}
__device__ __forceinline__ SF st_comb(const SF &a, const SF &b) { return st_combf<SF>(a, b); }
__device__ __forceinline__ SD st_comb(const SD &a, const SD &b) { return st_combf<SD>(a, b); }
__device__ __forceinline__ SI st_comb(const SI &a, const SI &b) {
  if (!b.n) return a; if (!a.n) return b;
  SI r; r.s = a.s + b.s; const i64 h = a.s + b.hi, l = a.s + b.lo;
  r.hi = a.hi > h ? a.hi : h; r.lo = a.lo < l ? a.lo : l; r.m = b.m > a.m ? b.m : a.m; r.n = 1; return r;
}
__device__ __forceinline__ SF st_shfl_down(const SF &a, int off) { SF r; r.s = __shfl_down(a.s, off, 64); r.c = __shfl_down(a.c, off, 64); r.m = __shfl_down(a.m, off, 64); return r; }
__device__ __forceinline__ SD st_shfl_down(const SD &a, int off) { SD r; r.s = __shfl_down(a.s, off, 64); r.c = __shfl_down(a.c, off, 64); r.m = __shfl_down(a.m, off, 64); return r; }
__device__ __forceinline__ SI st_shfl_down(const SI &a, int off) { SI r; r.s = __shfl_down(a.s, off, 64); r.hi = __shfl_down(a.hi, off, 64); r.lo = __shfl_down(a.lo, off, 64); r.m = __shfl_down(a.m, off, 64); r.n = __shfl_down(a.n, off, 64); return r; }
__device__ __forceinline__ double st_sum(const SF &a) { return (double)a.s + (double)a.c; }
__device__ __forceinline__ double st_sum(const SD &a) { return a.s + a.c; }
__device__ __forceinline__ double st_sum(const SI &a) { return (double)a.s; }
__device__ __forceinline__ double st_max(const SF &a) { return (double)a.m; }
__device__ __forceinline__ double st_max(const SD &a) { return a.m; }
__device__ __forceinline__ double st_max(const SI &a) { return (double)a.m; }
__device__ __forceinline__ int st_ovf(const SF &) { return 0; }
__device__ __forceinline__ int st_ovf(const SD &) { return 0; }
__device__ __forceinline__ int st_ovf(const SI &a) { return (a.n && (a.hi > 2147483647LL || a.lo < -2147483648LL)) ? 1 : 0; }
// Ordered fold of a wavefront: lane l ends up holding lanes [l, 64) folded in
// lane order; lane 0 holds the whole wavefront.
template <typename S> __device__ __forceinline__ S wave_fold(S v) {
  const int lane = threadIdx.x & 63;
  for (int off = 1; off < 64; off <<= 1) {
    const S o = st_shfl_down(v, off);
    if ((lane & (2 * off - 1)) == 0) v = st_comb(v, o);
  }
  return v;
}
struct RedArgs {
  const void *in[8]; i64 in_off[8];
  void *rowpart; void *colpart; void *rowfinal;
  double *full_sum, *full_max, *row_sum, *row_max, *col_sum, *col_max;
  int *flags;
  i64 *rowzero;
};
)PH";

// Host mirrors of the kernels' reduction states (kRedPrelude), for sizing
// the scratch that holds them.
struct HostSF { float s, c; double m; };
struct HostSD { double s, c, m; };
struct HostSI { int64_t s, hi, lo; int32_t m, n; };
static_assert(sizeof(HostSF) == 16 && sizeof(HostSD) == 24 && sizeof(HostSI) == 32, "state layouts");

struct RedArgs {
  const void *in[PH_MAX_INPUTS];
  int64_t in_off[PH_MAX_INPUTS];
  void *rowpart, *colpart, *rowfinal;
  double *full_sum, *full_max, *row_sum, *row_max, *col_sum, *col_max;
  int *flags;
  int64_t *rowzero;  // per block of 256 rows: its first zero's (view-order index << 1 | sign), or INT64_MAX
};

struct RedPlan {
  int64_t rows, cols;
  int tr;         // rows per tile
  int64_t ntr, ntc;
  int vec;        // columns per lane
  bool do_rows, do_cols, do_full;
  const char *state;   // SF / SD / SI
  size_t state_bytes;
  const char *acc_t;   // float / double / i64
  bool vec_ok[PH_MAX_INPUTS];  // whole-vector loads of the lane's VEC columns
  bool colex = false;  // the view's lexicographic order runs down the columns (a swapped transposed view)
  int64_t ntc_fin = 0;  // column-tile states per row left for the finish after the pre-folds
  int64_t ntr_fin = 0;  // row-tile states per column, likewise
};

// Column-tile states per row that the finish folds on one thread; rows with
// more are pre-folded in ordered runs of this many (ph_red_prefold), and so
// are the row-tile states of columns with more.
constexpr int64_t kRowFold = 64;
int64_t prefolded(int64_t ntc) {
  while (ntc > kRowFold) ntc = (ntc + kRowFold - 1) / kRowFold;
  return ntc;
}

// Statements evaluating the program at element (R, C) of the 2-D
// arrangement; returns the value as a double expression.
std::string emit_elem_at(const ph_prog *p, int n_in, const int *in_dt, const int64_t *rs, const int64_t *cs,
                         const std::string &R, const std::string &C, const std::string &sfx, std::string *code) {
  std::vector<std::string> inexpr;
  for (int j = 0; j < n_in; ++j) {
    const std::string x = fmt("e%d%s", j, sfx.c_str());
    *code += fmt("      const %s %s = ph_ld((const %s *)A.in[%d], A.in_off[%d] + (%s) * %s + (%s) * %s);\n",
                 ctype(compute_dt(in_dt[j])), x.c_str(), stype(in_dt[j]), j, j, R.c_str(), lit_i64(rs[j]).c_str(),
                 C.c_str(), lit_i64(cs[j]).c_str());
    inexpr.push_back(x);
  }
  const std::string res = emit_prog(p, inexpr, code, sfx);
  return "(double)(" + res + ")";
}

// Tile kernel for floating-point accumulation (states SF / SD).  A
// workgroup owns TR rows x 256*VEC columns; lane l of wavefront w owns VEC
// consecutive columns.  The value of every element is summed in FP64 twice:
// into the lane's column sums (one per column, running over the tile's rows)
// and into a per-row lane sum over its VEC columns.  Row sums of 8 rows at a
// time are folded across the wavefront by a transposing butterfly; the four
// wavefronts of the workgroup and, in the finish kernel, the tiles are then
// combined as compensated states with error-free TwoSum.  Maxima are taken in
// the value's own type (exact).  Rows are processed in groups of 8: the loads
// of the next group are issued before the current one is computed (two
// register buffers, alternating).  Bounds checks are emitted only when the
// shape is not a whole number of tiles, and then only in a second copy of
// the body that the last row and column tiles run; whole tiles run one free
// of predication.
std::string red_tile_float(const ph_prog *p, int n_in, const int *in_dt, const int64_t *rs,
                           const int64_t *cs, const RedPlan &P) {
  const int V = P.vec;
  const bool sd = std::string(P.state) == "SD";
  const char *M = (sd || p->out_dtype == PH_F64) ? "double" : "float";
  // The maxima take quiet operands only: a value that may be a signalling
  // NaN (a load passed through) is quieted once, by canonicalisation, where
  // it is computed (the compiler's max leaves the result for a signalling
  // operand unspecified, and a row's chain then lost its earlier values).
  // The column maxima are then taken by a v_max instruction in an asm
  // statement, which also keeps each column's chain in row order; the
  // compiler's own max would re-quiet a value coming out of an asm
  // statement, one extra instruction per pinned maximum.
  const bool quiet = max_operand_quiet(p, std::string(M) == "double" && p->out_dtype != PH_F64);
  const bool full_cols = P.cols % (256 * V) == 0;
  const bool full_rows = P.rows % P.tr == 0;
  const std::string C = lit_i64(P.cols);
  // bounds checks emitted by the generators below (switched off for the body
  // that whole tiles run)
  bool chk_cols = !full_cols, chk_rows = !full_rows;
  // Lane-granular column handling: when every input takes whole-vector loads
  // and the row length is a multiple of VEC, each lane's columns lie wholly
  // inside or wholly outside the row.  Lanes outside then load a clamped
  // (valid) vector, accumulate column values that are never written, and
  // give their rows zero and -inf: no per-element branches.
  bool lane_ok = P.cols % V == 0;
  for (int j = 0; j < n_in; ++j) lane_ok = lane_ok && P.vec_ok[j];
  bool lane_cols = false;
  // Column of a lane's v-th value: VEC consecutive columns when every input
  // takes whole-vector loads; otherwise columns 64 apart, so that each
  // element-wise load instruction reads 64 consecutive elements.
  bool all_vec = true;
  for (int j = 0; j < n_in; ++j) all_vec = all_vec && P.vec_ok[j];
  auto colv = [&](const std::string &v) { return all_vec ? "col0 + " + v : "cb + lane + 64 * (" + v + ")"; };
  auto ccheck = [&](const std::string &expr) { return chk_cols ? "(" + expr + " < " + C + ")" : std::string("true"); };
  auto rcheck = [&]() { return chk_rows ? std::string("(r < r1)") : std::string("true"); };
  // loads of the 8 rows starting at row `g` into buffer `b`
  // (rows past the tile's end read its last row again: every group's loads
  // are issued unconditionally, so no branch separates them from the wait for
  // the previous group's data)
  auto loads = [&](const std::string &b, const std::string &g) {
    std::string o;
    o += "#pragma unroll\n    for (int k = 0; k < 8; ++k) {\n      const i64 r = (" + g + " + k < r1) ? " + g + " + k : r1 - 1;\n";
    o += "      {\n";
    for (int j = 0; j < n_in; ++j) {
      o += fmt("        const %s *rp%d = (const %s *)A.in[%d] + (A.in_off[%d] + r * %s);\n", stype(in_dt[j]), j,
               stype(in_dt[j]), j, j, lit_i64(rs[j]).c_str());
      if (lane_cols) {
        o += fmt("        ph_ldv<%d>(rp%d + col0c, %s%d[k]);\n", V, j, b.c_str(), j);
        continue;
      }
      if (P.vec_ok[j] && !chk_cols) {
        // the row's address is uniform, the lane's byte offset 32-bit
        o += fmt("        ph_ldv<%d>((const %s *)((const char *)(rp%d + ct * %d) + lane_b%d), %s%d[k]);\n", V,
                 stype(in_dt[j]), j, 256 * V, j, b.c_str(), j);
        continue;
      }
      // element by element; with column checks a column past the row's end
      // reads the row's last element (its values are masked where used)
      const std::string CS = lit_i64(cs[j]);
      const std::string cc = chk_cols ? "(c < " + C + " ? c : " + lit_i64(P.cols - 1) + ")" : "c";
      if (in_dt[j] == PH_BF16 || in_dt[j] == PH_BOOL) {
        // into the packed words
        const int per = in_dt[j] == PH_BF16 ? 2 : 4, bits = in_dt[j] == PH_BF16 ? 16 : 8;
        o += fmt("#pragma unroll\n        for (int q = 0; q < %d; ++q) {\n          unsigned w = 0;\n", V / per);
        o += fmt("#pragma unroll\n          for (int e = 0; e < %d; ++e) {\n            const i64 c = %s;\n", per,
                 colv(fmt("q * %d + e", per)).c_str());
        o += fmt("            w |= (unsigned)rp%d[%s * %s] << (%d * e);\n", j, cc.c_str(), CS.c_str(), bits);
        o += fmt("          }\n          %s%d[k][q] = w;\n        }\n", b.c_str(), j);
        continue;
      }
      o += fmt("#pragma unroll\n        for (int v = 0; v < %d; ++v) {\n          const i64 c = %s;\n", V, colv("v").c_str());
      o += fmt("          %s%d[k][v] = ph_ld(rp%d, %s * %s);\n", b.c_str(), j, j, cc.c_str(), CS.c_str());
      o += "        }\n";
    }
    o += "      }\n    }\n";
    return o;
  };
  // the 8 rows at row `g` held in buffer `b`
  auto compute = [&](const std::string &b, const std::string &g) {
    std::vector<std::string> inexpr;
    for (int j = 0; j < n_in; ++j) {
      if (in_dt[j] == PH_BF16)
        inexpr.push_back(fmt("ph_bf(%s%d[k][v >> 1], v & 1)", b.c_str(), j));
      else if (in_dt[j] == PH_BOOL)
        inexpr.push_back(fmt("ph_bb(%s%d[k][v >> 2], v & 3)", b.c_str(), j));
      else
        inexpr.push_back(fmt("%s%d[k][v]", b.c_str(), j));
    }
    std::string body;
    const std::string res = emit_prog(p, inexpr, &body, "");
    std::string o = "    {\n";
    if (P.do_rows) o += "    double rsv[8]; M rmv[8];\n";
    o += "#pragma unroll\n    for (int k = 0; k < 8; ++k) {\n      const i64 r = " + g + " + k;\n";
    o += "      double ls = 0.0; M lm = -__builtin_inf();\n";
    o += "      if (" + rcheck() + ") {\n";
    o += fmt("#pragma unroll\n        for (int v = 0; v < %d; ++v) {\n          const i64 c = %s;\n", V, colv("v").c_str());
    o += "          {\n";
    o += body;
    o += fmt(quiet ? "            const double val = (double)(%s); const M mval = (M)(%s);\n"
                   : "            const double val = (double)(%s); const M mval = ph_quiet((M)(%s));\n",
             res.c_str(), res.c_str());
    // (columns past the row's end accumulate values that are never written)
    if (P.do_cols) o += "            csum[v] += val; ph_amax(cmax[v], mval);\n";
    if (P.do_rows) {
      // with every column present the lane's first column starts the row
      // (the row's state still starts from +0.0 where it is combined);
      // otherwise columns past the row's end are masked out of it
      if (!chk_cols)
        o += "            ls = v == 0 ? val : ls + val; lm = v == 0 ? mval : ph_fmax(lm, mval);\n";
      else
        o += "            const bool okc = " + ccheck("c") + ";\n            ls += okc ? val : 0.0; lm = okc ? ph_fmax(lm, mval) : lm;\n";
    }
    o += "          }\n        }\n      }\n";
    if (P.do_rows && lane_cols) o += "      ls = lane_in ? ls : 0.0; lm = lane_in ? lm : (M)-__builtin_inf();\n";
    if (P.do_rows) o += "      rsv[k] = ls; rmv[k] = lm;\n";
    if (P.do_cols) {
      // The column sums are brought up to date row by row: an empty asm
      // that takes the running value as in-out operand (the maxima are
      // updated by their own asm statement).  Left alone, the compiler
      // defers the column chains to the end of the group and holds every
      // element's value of 8 rows until then, which does not fit in
      // registers.
      o += fmt("#pragma unroll\n      for (int v = 0; v < %d; ++v) __asm__ volatile(\"\" : \"+v\"(csum[v]));\n", V);
    }
    o += "    }\n";
    if (P.do_rows) {
      o += "    const double fs = tfold8<double, false>(rsv);\n    const M fm = tfold8<M, true, true>(rmv);\n";
      o += "    const int rr = (lane >> 3) & 7;\n";
      o += "    if ((lane & 7) == 0 && " + g + " + rr < r1) { rsum_sh[" + g + " - r0 + rr][wave] = fs; rmax_sh[" + g +
           " - r0 + rr][wave] = fm; }\n";
    }
    o += "    }\n";
    return o;
  };
  std::string s;
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_red_tile(RedArgs A) {\n";
  s += fmt("  typedef %s S; typedef %s M;\n", P.state, M);
  s += fmt("  const i64 ct = blockIdx.x %% %s; const i64 rt = blockIdx.x / %s;\n", lit_i64(P.ntc).c_str(),
           lit_i64(P.ntc).c_str());
  s += "  const int lane = threadIdx.x & 63; const int wave = threadIdx.x >> 6;\n";
  s += fmt("  const i64 cb = ct * %d + wave * %d, col0 = cb + lane * %d;\n", 256 * V, 64 * V, V);
  s += fmt("  const i64 r0 = rt * %d;\n", P.tr);
  s += fmt("  const i64 r1 = (r0 + %d < %s) ? r0 + %d : %s;\n", P.tr, lit_i64(P.rows).c_str(), P.tr,
           lit_i64(P.rows).c_str());
  s += "  int fl = 0;\n";
  if (P.do_full) s += "  if (blockIdx.x == 0 && threadIdx.x == 0) A.flags[1] = 0;  // the finish kernel's ticket\n";
  s += fmt("  double csum[%d]; M cmax[%d];\n#pragma unroll\n  for (int v = 0; v < %d; ++v) { csum[v] = 0.0; cmax[v] = -__builtin_inf(); }\n",
           V, V, V);
  if (P.do_rows) s += fmt("  __shared__ double rsum_sh[%d][4]; __shared__ M rmax_sh[%d][4];\n", P.tr, P.tr);
  for (int j = 0; j < n_in; ++j) {
    if (in_dt[j] == PH_BF16 || in_dt[j] == PH_BOOL) {
      const int w = in_dt[j] == PH_BF16 ? V / 2 : V / 4;
      s += fmt("  unsigned xa%d[8][%d], xb%d[8][%d];\n", j, w, j, w);
    } else {
      const char *T = ctype(in_dt[j]);
      s += fmt("  %s xa%d[8][%d], xb%d[8][%d];\n", T, j, V, j, V);
    }
  }
  for (int j = 0; j < n_in; ++j)
    if (P.vec_ok[j])
      s += fmt("  const unsigned lane_b%d = threadIdx.x * %du;\n", j, V * (int)ph::dtype_size(in_dt[j]));
  auto main_loop = [&]() {
    std::string o = "  {\n" + loads("xa", "r0") + "  }\n";
    o += "  for (i64 g0 = r0; g0 < r1; g0 += 16) {\n";
    o += "    {\n" + loads("xb", "(g0 + 8)") + "    }\n";
    o += compute("xa", "g0");
    o += "    {\n" + loads("xa", "(g0 + 16)") + "    }\n";
    o += compute("xb", "(g0 + 8)");
    o += "  }\n";
    return o;
  };
  if (full_rows && full_cols) {
    s += main_loop();
  } else {
    // whole tiles (every row and column present: all but the last row and
    // column tiles) run a body without bounds checks
    s += fmt("  if (r1 - r0 == %d && ct < %s) {\n", P.tr, full_cols ? lit_i64(P.ntc).c_str() : lit_i64(P.ntc - 1).c_str());
    chk_rows = chk_cols = false;
    s += main_loop();
    s += "  } else {\n";
    chk_rows = !full_rows;
    chk_cols = !full_cols;
    if (chk_cols && lane_ok) {
      chk_cols = false;
      lane_cols = true;
      s += fmt("  const bool lane_in = col0 + %d <= %s;\n", V, C.c_str());
      s += fmt("  const i64 col0c = lane_in ? col0 : %s;\n", lit_i64(P.cols - V).c_str());
    }
    s += main_loop();
    lane_cols = false;
    chk_cols = !full_cols;
    s += "  }\n";
  }
  if (P.do_cols) {
    s += fmt("#pragma unroll\n  for (int v = 0; v < %d; ++v) {\n    const i64 c = %s;\n", V, colv("v").c_str());
    s += "    if (" + ccheck("c") + fmt(") { S st; st_from(st, csum[v], (double)cmax[v]); ((S *)A.colpart)[rt * %s + c] = st; }\n  }\n", C.c_str());
  }
  if (P.do_rows) {
    s += "  __syncthreads();\n";
    s += "  for (i64 i = threadIdx.x; i < r1 - r0; i += 256) {\n";
    s += "    S a; st_from(a, rsum_sh[i][0], (double)rmax_sh[i][0]);\n";
    s += "    for (int w = 1; w < 4; ++w) { S b; st_from(b, rsum_sh[i][w], (double)rmax_sh[i][w]); a = st_comb(a, b); }\n";
    s += fmt("    ((S *)A.rowpart)[ct * %s + r0 + i] = a;\n  }\n", lit_i64(P.rows).c_str());
  }
  s += "  if (fl) atomicOr(A.flags, fl);\n}\n";
  return s;
}

std::string red_source(const ph_prog *p, int n_in, const int *in_dt, const int64_t *rs,
                       const int64_t *cs, const RedPlan &P) {
  std::string s = kPrelude;
  s += kRedPrelude;
  const int V = P.vec;
  if (std::string(P.state) != "SI") {
    s += red_tile_float(p, n_in, in_dt, rs, cs, P);
  } else {
  // --- tile kernel (integer states: every fold in lexicographic order)
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_red_tile(RedArgs A) {\n";
  s += fmt("  typedef %s S;\n", P.state);
  s += fmt("  const i64 ct = blockIdx.x %% %s; const i64 rt = blockIdx.x / %s;\n", lit_i64(P.ntc).c_str(),
           lit_i64(P.ntc).c_str());
  s += "  const int lane = threadIdx.x & 63; const int wave = threadIdx.x >> 6;\n";
  s += fmt("  const i64 col0 = ct * %d + wave * %d + lane * %d;\n", 256 * V, 64 * V, V);
  s += fmt("  const i64 r0 = rt * %d;\n", P.tr);
  s += fmt("  const i64 r1 = (r0 + %d < %s) ? r0 + %d : %s;\n", P.tr, lit_i64(P.rows).c_str(), P.tr,
           lit_i64(P.rows).c_str());
  s += "  int fl = 0;\n";
  if (P.do_full) s += "  if (blockIdx.x == 0 && threadIdx.x == 0) A.flags[1] = 0;  // the finish kernel's ticket\n";
  s += fmt("  S cst[%d];\n#pragma unroll\n  for (int v = 0; v < %d; ++v) st_init(cst[v]);\n", V, V);
  if (P.do_rows) s += fmt("  __shared__ S rowsh[%d][4];\n", P.tr);
  s += "  for (i64 r = r0; r < r1; ++r) {\n";
  for (int j = 0; j < n_in; ++j)
    s += fmt("    const i64 rb%d = A.in_off[%d] + r * %s;\n", j, j, lit_i64(rs[j]).c_str());
  for (int j = 0; j < n_in; ++j) s += fmt("    %s x%d[%d];\n", ctype(compute_dt(in_dt[j])), j, V);
  for (int j = 0; j < n_in; ++j) {
    if (P.vec_ok[j]) {
      // whole-vector load when the lane's VEC columns are in range and aligned
      const char *vt = in_dt[j] == PH_F64 ? "double4" : in_dt[j] == PH_BF16 ? "ushort4"
                       : in_dt[j] == PH_I32 ? "int4" : in_dt[j] == PH_BOOL ? "uchar4" : "float4";
      s += fmt("    if (col0 + %d <= %s) {\n", V, lit_i64(P.cols).c_str());
      s += fmt("      const %s t = *(const %s *)((const %s *)A.in[%d] + rb%d + col0);\n", vt, vt,
               stype(in_dt[j]), j, j);
      s += fmt("      const %s tt[4] = {t.x, t.y, t.z, t.w};\n", stype(in_dt[j]));
      s += fmt("#pragma unroll\n      for (int v = 0; v < 4; ++v) x%d[v] = ph_ld(tt, v);\n", j);
      s += "    } else {\n";
    } else {
      s += "    {\n";
    }
    s += fmt("#pragma unroll\n      for (int v = 0; v < %d; ++v) {\n        const i64 c = col0 + v;\n", V);
    s += fmt("        if (c < %s) x%d[v] = ph_ld((const %s *)A.in[%d], rb%d + c * %s);\n", lit_i64(P.cols).c_str(),
             j, stype(in_dt[j]), j, j, lit_i64(cs[j]).c_str());
    s += "      }\n    }\n";
  }
  if (P.do_rows) s += "    S rst; st_init(rst);\n";
  std::vector<std::string> inexpr;
  for (int j = 0; j < n_in; ++j) inexpr.push_back(fmt("x%d[v]", j));
  std::string body;
  std::string res = emit_prog(p, inexpr, &body, "");
  s += fmt("#pragma unroll\n    for (int v = 0; v < %d; ++v) {\n      const i64 c = col0 + v;\n", V);
  s += fmt("      if (c < %s) {\n", lit_i64(P.cols).c_str());
  s += body;
  s += fmt("        const %s val = (%s)(%s);\n", P.acc_t, P.acc_t, res.c_str());
  if (P.do_cols) s += "        st_add(cst[v], val);\n";
  if (P.do_rows) s += "        st_add(rst, val);\n";
  s += "      }\n    }\n";
  if (P.do_rows) {
    s += "    rst = wave_fold(rst);\n";
    s += "    if (lane == 0) rowsh[r - r0][wave] = rst;\n";
  }
  s += "  }\n";
  if (P.do_cols) {
    s += fmt("#pragma unroll\n  for (int v = 0; v < %d; ++v) {\n    const i64 c = col0 + v;\n", V);
    s += fmt("    if (c < %s) ((S *)A.colpart)[rt * %s + c] = cst[v];\n  }\n", lit_i64(P.cols).c_str(),
             lit_i64(P.cols).c_str());
  }
  if (P.do_rows) {
    s += "  __syncthreads();\n";
    s += "  for (i64 i = threadIdx.x; i < r1 - r0; i += 256) {\n";
    s += "    S a = rowsh[i][0]; a = st_comb(a, rowsh[i][1]); a = st_comb(a, rowsh[i][2]); a = st_comb(a, rowsh[i][3]);\n";
    s += fmt("    ((S *)A.rowpart)[ct * %s + r0 + i] = a;\n  }\n", lit_i64(P.rows).c_str());
  }
  s += "  if (fl) atomicOr(A.flags, fl);\n}\n";
  }
  if ((P.do_rows && P.ntc > kRowFold) || (P.do_cols && P.ntr > kRowFold)) {
    // pre-fold: state t of series i (a row over its column tiles, or a column
    // over its row tiles) <- the states [t * 64, t * 64 + 64) of the level
    // below, in order (a long series would otherwise be folded on one or four
    // threads of the finish)
    s += "struct FoldArgs { const void *src; void *dst; i64 rows; i64 nin; i64 nout; };\n";
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_red_prefold(FoldArgs F) {\n";
    s += fmt("  typedef %s S;\n", P.state);
    s += "  const i64 g = (i64)blockIdx.x * 256 + threadIdx.x;\n  if (g >= F.rows * F.nout) return;\n";
    s += "  const i64 i = g % F.rows, t = g / F.rows;\n";
    s += fmt("  const i64 t0 = t * %s, t1 = (t0 + %s < F.nin) ? t0 + %s : F.nin;\n", lit_i64(kRowFold).c_str(),
             lit_i64(kRowFold).c_str(), lit_i64(kRowFold).c_str());
    s += "  S a; st_init(a);\n#pragma unroll 8\n  for (i64 u = t0; u < t1; ++u) a = st_comb(a, ((const S *)F.src)[u * F.rows + i]);\n";
    s += "  ((S *)F.dst)[t * F.rows + i] = a;\n}\n";
  }
  // --- finish.  Column blocks: 64 columns x 4 parts; part p folds row tiles
  // [p*ntr/4, (p+1)*ntr/4) in order, then the 4 parts are folded in order.
  // Row blocks: one row per thread (its column tiles in order), then the
  // block's 256 rows are folded in order (a pairwise tree of adjacent
  // ranges) into one state per block for the full result.
  const int64_t ncb = P.do_cols ? (P.cols + 63) / 64 : 0;
  const bool fix = std::string(P.state) != "SI";  // float maxima follow Enumerable#max's `>`
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_red_finish(RedArgs A) {\n";
  s += fmt("  typedef %s S;\n", P.state);
  s += "  __shared__ S sh[256];\n  int fl = 0;\n";
  if (P.do_cols) {
    s += fmt("  if ((i64)blockIdx.x < %s) {\n", lit_i64(ncb).c_str());
    s += "    const int lane = threadIdx.x & 63; const int part = threadIdx.x >> 6;\n";
    s += "    const i64 c = (i64)blockIdx.x * 64 + lane;\n";
    s += fmt("    const i64 t0 = part * %s / 4, t1 = (part + 1) * %s / 4;\n", lit_i64(P.ntr_fin).c_str(),
             lit_i64(P.ntr_fin).c_str());
    s += "    S a; st_init(a);\n";
    if (fix) {
      // the column's first element, loaded before the fold so its latency overlaps it
      s += fmt("    double f0 = 0.0;\n    if (part == 0 && c < %s && A.col_max) {\n", lit_i64(P.cols).c_str());
      const std::string e_ = emit_elem_at(p, n_in, in_dt, rs, cs, "0", "c", "_cf", &s);
      s += "      f0 = " + e_ + ";\n    }\n";
    }
    // (the states are loaded eight at a time before they are folded in
    // order, so the loads of a group are in flight together)
    const int64_t per = (P.ntr_fin + 3) / 4;
    s += fmt("    if (c < %s) {\n#pragma unroll\n      for (int g = 0; g < %s; g += 8) {\n        S v[8];\n"
             "#pragma unroll\n        for (int u = 0; u < 8; ++u) {\n"
             "          const i64 t = t0 + g + u;\n          if (g + u < %s && t < t1) v[u] = ((const S *)A.colpart)[t * %s + c];\n        }\n"
             "#pragma unroll\n        for (int u = 0; u < 8; ++u) if (g + u < %s && t0 + g + u < t1) a = st_comb(a, v[u]);\n      }\n    }\n",
             lit_i64(P.cols).c_str(), lit_i64(per).c_str(), lit_i64(per).c_str(), lit_i64(P.cols).c_str(),
             lit_i64(per).c_str());
    s += "    sh[threadIdx.x] = a;\n    __syncthreads();\n";
    if (fix) {
      // Enumerable#max keeps the slice's first element unless a later one is
      // greater: a leading NaN stays, and of equal zeros the first one's
      // sign.  Columns whose maximum is a zero are scanned by the whole
      // wavefront in turn, 64 rows at a time, for their first zero.
      s += "    if (part == 0) {\n      double mx = 0.0;\n      bool mzero = false;\n";
      s += fmt("      if (c < %s) {\n", lit_i64(P.cols).c_str());
      s += "        S r = sh[lane];\n        for (int q = 1; q < 4; ++q) r = st_comb(r, sh[q * 64 + lane]);\n";
      s += "        if (A.col_sum) { A.col_sum[c] = st_sum(r); fl |= st_ovf(r); }\n";
      s += "        mx = st_max(r);\n        mzero = A.col_max && f0 == f0 && mx == 0.0;\n      }\n";
      s += "      unsigned long long need = __ballot(mzero);\n";
      s += "      while (need) {\n        const int l = __builtin_ctzll(need);\n        need &= need - 1;\n";
      s += "        const i64 cl = (i64)blockIdx.x * 64 + l;\n";
      s += fmt("        bool found = false;\n        for (i64 r1 = 0; r1 < %s && !found; r1 += 512) {\n          double g[8];\n"
               "#pragma unroll\n          for (int k = 0; k < 8; ++k) {\n            const i64 rr = r1 + 64 * k + lane;\n"
               "            g[k] = 1.0;\n            if (rr < %s) {\n", lit_i64(P.rows).c_str(), lit_i64(P.rows).c_str());
      {
        std::string ev;
        const std::string e_ = emit_elem_at(p, n_in, in_dt, rs, cs, "rr", "cl", "_cz", &ev);
        s += ev + "              g[k] = " + e_ + ";\n            }\n          }\n";
      }
      s += "#pragma unroll\n          for (int k = 0; k < 8; ++k) {\n";
      s += "            const unsigned long long hit = __ballot(g[k] == 0.0);\n            if (hit && !found) {\n";
      s += "              const double gz = __shfl(g[k], __builtin_ctzll(hit));\n";
      s += "              if (lane == l) mx = gz;\n              found = true;\n            }\n          }\n        }\n      }\n";
      s += fmt("      if (c < %s && A.col_max) {\n        if (f0 != f0) mx = f0;\n        A.col_max[c] = mx;\n      }\n",
               lit_i64(P.cols).c_str());
    } else {
      s += fmt("    if (part == 0 && c < %s) {\n", lit_i64(P.cols).c_str());
      s += "      S r = sh[lane];\n      for (int q = 1; q < 4; ++q) r = st_comb(r, sh[q * 64 + lane]);\n";
      s += "      if (A.col_sum) { A.col_sum[c] = st_sum(r); fl |= st_ovf(r); }\n";
      s += "      if (A.col_max) A.col_max[c] = st_max(r);\n";
    }
    s += "    }\n  } else {\n";
  } else {
    s += "  {\n";
  }
  if (P.do_rows) {
    s += fmt("    const i64 rb = (i64)blockIdx.x - %s;\n", lit_i64(ncb).c_str());
    s += "    const i64 i = rb * 256 + threadIdx.x;\n    S a; st_init(a);\n";
    if (fix) {
      // the row's first element, loaded before the fold so its latency overlaps it
      s += fmt("    double f0 = 0.0;\n    if (i < %s) {\n", lit_i64(P.rows).c_str());
      const std::string e_ = emit_elem_at(p, n_in, in_dt, rs, cs, "i", "0", "_rf", &s);
      s += "      f0 = " + e_ + ";\n    }\n";
      s += "    i64 zkey = 0x7fffffffffffffffLL;\n";
      if (P.do_full) {
        // the view's first element for the full maximum, likewise loaded early
        // (only the row block that finishes last uses it)
        s += "    double ff0 = 0.0;\n    if (threadIdx.x == 0 && A.full_max) {\n";
        const std::string e2 = emit_elem_at(p, n_in, in_dt, rs, cs, "0", "0", "_ff", &s);
        s += "      ff0 = " + e2 + ";\n    }\n";
      }
    }
    if (fix) s += "    double mx = 0.0;\n    bool mzero = false;\n";
    s += fmt("    if (i < %s) {\n#pragma unroll\n      for (int g = 0; g < %s; g += 8) {\n        S v[8];\n"
             "#pragma unroll\n        for (int u = 0; u < 8; ++u) if (g + u < %s) v[u] = ((const S *)A.rowpart)[(g + u) * %s + i];\n"
             "#pragma unroll\n        for (int u = 0; u < 8; ++u) if (g + u < %s) a = st_comb(a, v[u]);\n      }\n",
             lit_i64(P.rows).c_str(), lit_i64(P.ntc_fin).c_str(), lit_i64(P.ntc_fin).c_str(), lit_i64(P.rows).c_str(),
             lit_i64(P.ntc_fin).c_str());
    s += "      if (A.row_sum) { A.row_sum[i] = st_sum(a); fl |= st_ovf(a); }\n";
    if (fix) {
      // first element and, when the row's maximum is a zero, the first zero
      // (its view-order index and sign also serve the full result)
      s += "      mx = st_max(a);\n      mzero = mx == 0.0;\n    }\n";
      // Rows of this wavefront whose maximum is a zero: the wavefront scans
      // each in turn, 64 consecutive columns at a time, and the lowest lane
      // holding a zero has the row's first one.  (A zero maximum is one of
      // the row's values, so the scan ends within the row.)
      s += "    {\n      const int lane = threadIdx.x & 63;\n      unsigned long long need = __ballot(mzero);\n";
      s += "      while (need) {\n        const int l = __builtin_ctzll(need);\n        need &= need - 1;\n";
      s += "        const i64 ri = i - lane + l;\n";
      // (eight chunks of 64 columns are loaded per round, then checked in order)
      s += fmt("        bool found = false;\n        for (i64 c1 = 0; c1 < %s && !found; c1 += 512) {\n          double g[8];\n"
               "#pragma unroll\n          for (int k = 0; k < 8; ++k) {\n            const i64 cc = c1 + 64 * k + lane;\n"
               "            g[k] = 1.0;\n            if (cc < %s) {\n", lit_i64(P.cols).c_str(), lit_i64(P.cols).c_str());
      {
        std::string ev;
        const std::string e_ = emit_elem_at(p, n_in, in_dt, rs, cs, "ri", "cc", "_rz", &ev);
        s += ev + "              g[k] = " + e_ + ";\n            }\n          }\n";
      }
      const std::string key =
          P.colex ? fmt("zc * %s + ri", lit_i64(P.rows).c_str()) : fmt("ri * %s + zc", lit_i64(P.cols).c_str());
      s += "#pragma unroll\n          for (int k = 0; k < 8; ++k) {\n";
      s += "            const unsigned long long hit = __ballot(g[k] == 0.0);\n            if (hit && !found) {\n";
      s += "              const int hl = __builtin_ctzll(hit);\n              const double gz = __shfl(g[k], hl);\n";
      s += "              const i64 zc = c1 + 64 * k + hl;\n";
      s += "              if (lane == l) { mx = gz; zkey = ((" + key + ") << 1) | (__builtin_signbit(gz) ? 1 : 0); }\n";
      s += "              found = true;\n            }\n          }\n        }\n      }\n    }\n";
      s += fmt("    if (i < %s) {\n      if (f0 != f0) mx = f0;\n      if (A.row_max) A.row_max[i] = mx;\n    }\n",
               lit_i64(P.rows).c_str());
    } else {
      s += "      if (A.row_max) A.row_max[i] = st_max(a);\n    }\n";
    }
    if (P.do_full) {
      const int64_t nrb = (P.rows + 255) / 256;
      // (with the fix-up, the block's least first-zero key is folded alongside)
      if (fix) s += "    __shared__ i64 zsh[256];\n    zsh[threadIdx.x] = zkey;\n";
      s += "    sh[threadIdx.x] = a;\n    __syncthreads();\n";
      s += "    for (int off = 1; off < 256; off <<= 1) {\n      S o = sh[threadIdx.x];\n";
      if (fix) s += "      i64 zo = zsh[threadIdx.x];\n";
      s += "      if ((threadIdx.x & (2 * off - 1)) == 0) {\n        o = st_comb(o, sh[threadIdx.x + off]);\n";
      if (fix) s += "        zo = zsh[threadIdx.x + off] < zo ? zsh[threadIdx.x + off] : zo;\n";
      s += "      }\n      __syncthreads();\n      sh[threadIdx.x] = o;\n";
      if (fix) s += "      zsh[threadIdx.x] = zo;\n";
      s += "      __syncthreads();\n    }\n";
      // The row block that finishes last (a ticket counter, reset by the
      // tile kernel) folds the row blocks' states in order for the full
      // result: one wavefront, each lane a run of consecutive blocks, then
      // an ordered fold across the lanes.
      s += "    __shared__ int last;\n";
      s += "    if (threadIdx.x == 0) {\n      ((S *)A.rowfinal)[rb] = sh[0];\n";
      if (fix) s += "      A.rowzero[rb] = zsh[0];\n";
      s += "      __threadfence();\n";
      s += fmt("      last = atomicAdd(A.flags + 1, 1) == %s - 1;\n    }\n    __syncthreads();\n", lit_i64(nrb).c_str());
      s += "    if (last && threadIdx.x < 64) {\n      __threadfence();\n";
      s += fmt("      const i64 n = %s; const i64 per = (n + 63) / 64;\n", lit_i64(nrb).c_str());
      s += "      const i64 b = threadIdx.x * per; const i64 e = (b + per < n) ? b + per : n;\n";
      s += "      S f; st_init(f);\n      for (i64 j = b; j < e; ++j) f = st_comb(f, ((const S *)A.rowfinal)[j]);\n";
      s += "      f = wave_fold(f);\n";
      if (fix) {
        // the first zero of the whole view: the least key over the row blocks
        // (after wave_fold only lane 0 holds the whole maximum)
        s += "      i64 zk = 0x7fffffffffffffffLL;\n      const double fm0 = __shfl(st_max(f), 0);\n";
        s += "      if (A.full_max && fm0 == 0.0)\n        for (i64 j = threadIdx.x; j < " +
             lit_i64(nrb) + "; j += 64) zk = A.rowzero[j] < zk ? A.rowzero[j] : zk;\n";
        s += "      for (int o = 32; o > 0; o >>= 1) { const i64 t = __shfl_xor(zk, o); zk = t < zk ? t : zk; }\n";
      }
      s += "      if (threadIdx.x == 0) {\n        if (A.full_sum) { A.full_sum[0] = st_sum(f); fl |= st_ovf(f); }\n";
      if (fix) {
        s += "        if (A.full_max) {\n          double mx = st_max(f);\n";
        s += "          if (ff0 != ff0) mx = ff0;\n";
        s += "          else if (mx == 0.0 && zk != 0x7fffffffffffffffLL) mx = (zk & 1) ? -0.0 : 0.0;\n";
        s += "          A.full_max[0] = mx;\n        }\n      }\n    }\n";
      } else {
        s += "        if (A.full_max) A.full_max[0] = st_max(f);\n      }\n    }\n";
      }
    }
  }
  s += "  }\n  if (fl) atomicOr(A.flags, fl);\n}\n";
  return s;
}

// Columns per lane of the reduction tile: 8 when every input is 2 bytes or
// narrower (16-byte BFloat16 loads), else 4; the integer tile always takes 4.
int red_vec(bool int_state, int n_in, const int *in_dt) {
  if (int_state || n_in < 1) return 4;
  for (int j = 0; j < n_in; ++j)
    if (ph::dtype_size(in_dt[j]) > 2) return 4;
  return 8;
}

// A 2-D view of every input: rows x cols with per-input row/col strides.
struct View2D {
  int64_t rows, cols;
  int64_t rs[PH_MAX_INPUTS], cs[PH_MAX_INPUTS], off[PH_MAX_INPUTS];
};

// Runs the fused reduction over a 2-D arrangement (inputs affine).
// A split of n elements into rows of `cols` that the tiles stream well: a
// multiple of 8 dividing n, nearest 16384 within [4096, 65536] (then
// [1024, 2^20]); 0 when n has none.
int64_t stream_split(int64_t n) {
  int64_t best = 0;
  for (int64_t lim : {int64_t(1) << 16, int64_t(1) << 20}) {
    const int64_t lo = lim == (int64_t(1) << 16) ? 4096 : 1024;
    for (int64_t c = lo; c <= lim && c <= n; c += 8)
      if (n % c == 0 && (best == 0 || std::llabs(c - 16384) < std::llabs(best - 16384))) best = c;
    if (best) break;
  }
  return best;
}

int run_reduce2d(const ph_prog *prog, int n_in, const void *const *ins, const int *in_dt,
                 const View2D &V0, int acc, double *full_sum, double *full_max, double *row_sum,
                 double *row_max, double *col_sum, double *col_max, int32_t *flags_out, bool colex) {
  View2D V = V0;
  // Only the whole view's result, over elements that lie contiguously in
  // lexicographic order (a rank-1 array, a whole row-major array): any
  // rows x cols split of the same run reads the same sequence, so the one
  // that streams best is taken (tall tiles, short rows).  Int32 running
  // totals and the maximum's first-zero key see the elements in the same
  // order; only the grouping of the float partial sums changes.
  if ((full_sum || full_max) && !row_sum && !row_max && !col_sum && !col_max && !colex) {
    bool contig = true;
    for (int j = 0; j < n_in; ++j) {
      // element (r, c) at off + r * rs + c * cs must be off + r * cols + c
      const bool cols_ok = V.cols == 1 || V.cs[j] == 1;
      const bool rows_ok = V.rows == 1 || V.rs[j] == V.cols;
      contig = contig && cols_ok && rows_ok;
    }
    const int64_t n = V.rows * V.cols;
    const int64_t c = contig ? stream_split(n) : 0;
    if (c && c != V.cols) {
      V.rows = n / c;
      V.cols = c;
      for (int j = 0; j < n_in; ++j) {
        V.rs[j] = c;
        V.cs[j] = 1;
      }
    }
  }
  RedPlan P;
  P.colex = colex;
  P.rows = V.rows;
  P.cols = V.cols;
  P.vec = 4;
  P.do_cols = col_sum || col_max;
  P.do_full = full_sum || full_max;
  P.do_rows = row_sum || row_max || P.do_full;
  const int odt = prog->out_dtype;
  if (odt == PH_I32 || odt == PH_BOOL) {
    P.state = "SI"; P.state_bytes = sizeof(HostSI); P.acc_t = "i64";
  } else if (acc == PH_ACC_F64) {
    P.state = "SD"; P.state_bytes = sizeof(HostSD); P.acc_t = "double";
  } else {
    P.state = "SF"; P.state_bytes = sizeof(HostSF); P.acc_t = "float";
  }
  const bool int_state = std::string(P.state) == "SI";
  P.vec = red_vec(int_state, n_in, in_dt);
  for (int j = 0; j < PH_MAX_INPUTS; ++j) P.vec_ok[j] = false;
  for (int j = 0; j < n_in; ++j) {
    // the float tile loads at most 16 bytes at a time, the integer tile one
    // 4-element vector
    const size_t vbytes = (size_t)P.vec * ph::dtype_size(in_dt[j]);
    const size_t align = int_state ? vbytes : std::min<size_t>(vbytes, 16);
    P.vec_ok[j] = V.cs[j] == 1 && V.off[j] % P.vec == 0 && (P.rows == 1 || V.rs[j] % P.vec == 0) &&
                  ((uintptr_t)ins[j] % align) == 0;
  }
  P.ntc = (P.cols + 256 * P.vec - 1) / (256 * P.vec);
  P.ntc_fin = prefolded(P.ntc);
  // rows per tile: aim for ~512 workgroups (fewer, taller tiles write fewer
  // column partial states; measured best on 16384^2), a multiple of 16 (the
  // float tile folds rows 8 at a time), at most 1024 rows (256 for integer
  // states, whose per-row LDS states are larger)
  int64_t tr = P.rows * P.ntc / 512;
  if (const char *e = std::getenv("PH_RED_TR")) tr = std::atol(e);  // tuning override
  const int64_t tr_cap = int_state ? 256 : 1024;
  if (tr > tr_cap) tr = tr_cap;
  tr = (tr + 15) / 16 * 16;
  if (tr < 16) tr = 16;
  if (int_state && tr > P.rows) tr = P.rows;
  if (tr < 1) tr = 1;
  P.tr = (int)tr;
  P.ntr = (P.rows + P.tr - 1) / P.tr;
  P.ntr_fin = prefolded(P.ntr);
  std::string key = "R" + prog_key(prog);
  key_add(&key, in_dt, sizeof(int) * (size_t)n_in);
  key_add(&key, V.rs, sizeof(int64_t) * (size_t)n_in);
  key_add(&key, V.cs, sizeof(int64_t) * (size_t)n_in);
  key_add(&key, P.vec_ok, sizeof(bool) * (size_t)n_in);
  const int64_t plan[] = {P.rows, P.cols, P.tr, P.ntr, P.ntc, P.vec, P.do_rows, P.do_cols, P.do_full, P.colex};
  key_add(&key, plan, sizeof(plan));
  key += P.state;
  hipFunction_t ktile, kfin, kfold = nullptr;
  const bool fold_rows = P.do_rows && P.ntc > kRowFold, fold_cols = P.do_cols && P.ntr > kRowFold;
  const bool prefold = fold_rows || fold_cols;
  int rc;
  auto hit = fused_cache().find(key);
  if (hit != fused_cache().end()) {
    ktile = hit->second.k[0];
    kfin = hit->second.k[1];
    kfold = hit->second.k[2];
  } else {
    std::string src = red_source(prog, n_in, in_dt, V.rs, V.cs, P);
    rc = ph::jit_kernel(src, "ph_red_tile", &ktile);
    if (rc) return rc;
    rc = ph::jit_kernel(src, "ph_red_finish", &kfin);
    if (rc) return rc;
    if (prefold) {
      rc = ph::jit_kernel(src, "ph_red_prefold", &kfold);
      if (rc) return rc;
    }
    fused_cache()[key] = FusedKernels{{ktile, kfin, kfold}, 0};
  }
  const bool may_flag = prog_may_flag(prog) || (std::string(P.state) == "SI" && (full_sum || row_sum || col_sum));
  // scratch: rowpart [ntc][rows], colpart [ntr][cols], rowfinal [row blocks of 256] (sized [rows]),
  // the first pre-fold levels [ntc / 64][rows] and [ntr / 64][cols]
  const size_t rowpart_b = P.do_rows ? (size_t)(P.ntc * P.rows) * P.state_bytes : 0;
  const size_t fold_b = fold_rows ? (size_t)((P.ntc + kRowFold - 1) / kRowFold * P.rows) * P.state_bytes : 0;
  const size_t cfold_b = fold_cols ? (size_t)((P.ntr + kRowFold - 1) / kRowFold * P.cols) * P.state_bytes : 0;
  const size_t colpart_b = P.do_cols ? (size_t)(P.ntr * P.cols) * P.state_bytes : 0;
  const size_t rowfinal_b = P.do_rows ? (size_t)P.rows * P.state_bytes : 0;
  const size_t rowzero_b = P.do_full ? (size_t)((P.rows + 255) / 256) * sizeof(int64_t) : 0;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  void *base;
  rc = ph::scratch(al(rowpart_b) + al(colpart_b) + al(rowfinal_b) + al(rowzero_b) + al(fold_b) + al(cfold_b) + 256,
                   &base, 2);
  if (rc) return rc;
  int *dflags;
  rc = get_flags_buffer(&dflags, may_flag);
  if (rc) return rc;
  RedArgs A;
  std::memset(&A, 0, sizeof(A));
  for (int j = 0; j < n_in; ++j) {
    A.in[j] = ins[j];
    A.in_off[j] = V.off[j];
  }
  char *b = (char *)base;
  A.rowpart = b;
  A.colpart = b + al(rowpart_b);
  A.rowfinal = b + al(rowpart_b) + al(colpart_b);
  A.rowzero = (int64_t *)(b + al(rowpart_b) + al(colpart_b) + al(rowfinal_b));
  A.full_sum = full_sum;
  A.full_max = full_max;
  A.row_sum = row_sum;
  A.row_max = row_max;
  A.col_sum = col_sum;
  A.col_max = col_max;
  A.flags = dflags;
  size_t asz = sizeof(A);
  void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &asz,
                 HIP_LAUNCH_PARAM_END};
  rc = ph::launch_gate("reduce");
  if (rc) return rc;
  PH_HIP(hipModuleLaunchKernel(ktile, (unsigned)(P.ntr * P.ntc), 1, 1, 256, 1, 1, 0, ph::stream(),
                               nullptr, cfg));
  // ordered pre-folds of long series of tile states, 64 to 1, ping-pong
  // between the tile kernel's buffer and a fold buffer, until the finish's
  // count is left; returns the buffer holding the last level
  auto fold = [&](void *states, char *spare, int64_t series, int64_t n) -> void * {
    struct {
      const void *src;
      void *dst;
      int64_t rows, nin, nout;
    } F;
    char *lv[2] = {(char *)states, spare};
    int cur = 0;
    while (n > kRowFold) {
      const int64_t nout = (n + kRowFold - 1) / kRowFold;
      F.src = lv[cur];
      F.dst = lv[cur ^ 1];
      F.rows = series;
      F.nin = n;
      F.nout = nout;
      size_t fsz = sizeof(F);
      void *fcfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &F, HIP_LAUNCH_PARAM_BUFFER_SIZE, &fsz, HIP_LAUNCH_PARAM_END};
      if (hipModuleLaunchKernel(kfold, (unsigned)((series * nout + 255) / 256), 1, 1, 256, 1, 1, 0, ph::stream(),
                                nullptr, fcfg) != hipSuccess)
        return nullptr;
      n = nout;
      cur ^= 1;
    }
    return lv[cur];
  };
  char *spare = (char *)A.rowzero + al(rowzero_b);
  if (fold_rows && !(A.rowpart = fold(A.rowpart, spare, P.rows, P.ntc)))
    return ph::set_error(PH_E_HIP, "reduction pre-fold launch failed");
  if (fold_cols && !(A.colpart = fold(A.colpart, spare + al(fold_b), P.cols, P.ntr)))
    return ph::set_error(PH_E_HIP, "reduction pre-fold launch failed");
  // finish: column blocks of 64 columns, then row blocks of 256 rows
  const int64_t nfin = (P.do_cols ? (P.cols + 63) / 64 : 0) + (P.do_rows ? (P.rows + 255) / 256 : 0);
  if (nfin > 0)
    PH_HIP(hipModuleLaunchKernel(kfin, (unsigned)nfin, 1, 1, 256, 1, 1, 0, ph::stream(), nullptr, cfg));
  return read_flags(dflags, may_flag, flags_out);
}

int check_inputs(const ph_prog *prog, int32_t n_in, const void *const *ins,
                 const ph_view_desc *in_descs, int *in_dt, bool allow_index = false) {
  if (n_in < 1 || n_in > PH_MAX_INPUTS)
    return ph::set_error(PH_E_ARGUMENT, "%d inputs (1..%d)", n_in, PH_MAX_INPUTS);
  for (int j = 0; j < n_in; ++j) {
    int rc = ph::validate_desc(&in_descs[j], "input");
    if (rc) return rc;
    if (!ins[j]) return ph::set_error(PH_E_ARGUMENT, "input %d: null pointer", j);
    in_dt[j] = in_descs[j].dtype;
    if (in_descs[j].view.rank != in_descs[0].view.rank)
      return ph::set_error(PH_E_SHAPE, "input %d rank differs", j);
    for (int i = 0; i < in_descs[0].view.rank; ++i)
      if (in_descs[j].view.shape[i] != in_descs[0].view.shape[i])
        return ph::set_error(PH_E_SHAPE, "input %d shape differs on axis %d", j, i);
  }
  return validate_prog(prog, n_in, in_dt, allow_index);
}

// Runs map_tiled_source's kernel (compiled once per program and layout).
int map_tiled(const ph_prog *prog, int n_in, const void *const *ins, const int *in_dt, const ph::Layout &L, int ax,
              unsigned tmask, void *out, const ph_view_desc *out_desc, int32_t *flags_out) {
  std::string key = "MT" + prog_key(prog);
  key_add(&key, in_dt, sizeof(int) * (size_t)n_in);
  key_add(&key, &L.rank, sizeof(L.rank));
  key_add(&key, L.shape, sizeof(int64_t) * (size_t)L.rank);
  for (int o = 0; o <= n_in; ++o) key_add(&key, L.stride[o], sizeof(int64_t) * (size_t)L.rank);
  const int32_t mk[3] = {out_desc->dtype, ax, (int32_t)tmask};
  key_add(&key, mk, sizeof(mk));
  hipFunction_t fn;
  int64_t nblocks = 0;
  auto hit = fused_cache().find(key);
  if (hit != fused_cache().end()) {
    fn = hit->second.k[0];
    nblocks = hit->second.nblocks;
  } else {
    std::string src = map_tiled_source(prog, n_in, in_dt, L, out_desc->dtype, ax, tmask, &nblocks);
    int rc = ph::jit_kernel(src, "ph_map_tiled", &fn);
    if (rc) return rc;
    fused_cache()[key] = FusedKernels{{fn, nullptr, nullptr}, nblocks};
  }
  const bool may_flag = prog_may_flag(prog);
  int *dflags;
  int rc = get_flags_buffer(&dflags, may_flag);
  if (rc) return rc;
  MapArgs A;
  std::memset(&A, 0, sizeof(A));
  for (int j = 0; j < n_in; ++j) {
    A.in[j] = ins[j];
    A.in_off[j] = L.offset[j];
  }
  A.out = out;
  A.out_off = L.offset[n_in];
  A.flags = dflags;
  size_t asz = sizeof(A);
  void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &asz, HIP_LAUNCH_PARAM_END};
  rc = ph::launch_gate("map (tiled)");
  if (rc) return rc;
  PH_HIP(hipModuleLaunchKernel(fn, (unsigned)grid_for(nblocks), 1, 1, kBlock, 1, 1, 0, ph::stream(), nullptr, cfg));
  return read_flags(dflags, may_flag, flags_out);
}

// ---------------------------------------------------------------------------
// Inputs that share the output's buffer
// ---------------------------------------------------------------------------
// The reference produces the elements of a map (or a copy) one at a time in
// lexicographic order of the output's shape and reads each element's inputs
// when it gets there (src/n_array.cr:484-492 pulls `src.each` lazily,
// :523-536 calls `mask.get` and `value.get` per coordinate).  An input that
// reads the buffer being written therefore sees the writes of the elements
// before it.  With the output map W injective, element L (its lexicographic
// index) reading address a of the shared buffer depends on element
// L' = W^-1(a) when that exists and L' < L; otherwise it reads the old value
// (L' > L is written later; L' = L is its own element, read before written).
//
// 1. A classifying pass finds whether any element reads an address another
//    element writes (a conflict) and whether any reads one written before it
//    (a read after write).  Without conflicts the ordinary parallel kernels
//    give the sequential result, in place.
// 2. Conflicts but no read after write: every read sees the old value; the
//    results go to scratch T, indexed by L, and are then written out.
// 3. Reads after writes: Jacobi iteration on T.  T_0 reads old values
//    everywhere; T_{m+1}(L) reads T_m(L') where L' < L and old values
//    elsewhere.  Dependencies point only to smaller L, so the fixed point is
//    unique and is the sequential result, and an element whose dependency
//    chain is d deep is exact from T_d on.  A pass that reproduces T bit for
//    bit has reached it, and its flags are the sequential run's.  Past
//    kOrdMaxIter passes (kOrdMaxIterWalk with one input on the output's
//    buffer; a deep chain, such as a shift by one element):
//    with one input on the output's buffer, every element depends on at
//    most one earlier element, and when none has two dependents the
//    dependencies are chains; a thread per chain root then runs its chain
//    in order, carrying each result to the next element (a recurrence down
//    the rows of a matrix runs a thread per column).  When that input reads
//    each address at most once through an affine view, its chains follow
//    in closed form (ph_ord_walkc) and the passes before are skipped.  When some element has
//    several dependents the dependencies form a tree.  If the elements that
//    others read still form chains, those chains are walked the same way
//    and one pass then computes the rest from their parents' results (a
//    broadcast of column 0 down the rows walks column 0).  Otherwise
//    pointer jumping finds each element's depth, the elements are ordered
//    by depth, and a pass per level evaluates them, each reading its
//    parent's result from T.  Otherwise
//    (several inputs on the buffer, or a tree too deep for passes per level)
//    one device thread runs the loop in order, kSeqChunk elements per launch.
// A plain copy takes pointer jumping instead of step 3: element L's value is
// the old value at the read address of its chain's root, and
// ceil(log2(depth)) passes over the chain links find every root.
constexpr int kOrdMaxIter = 12;
// With one input on the output's buffer the chain walk and the tree passes
// are exact at any depth, so the passes stop sooner: three still settle a map
// whose elements depend at most on elements that depend on none (a
// transposed self-assignment), and a deep chain pays 3 passes before its
// walk, not 13.
constexpr int kOrdMaxIterWalk = 2;
constexpr int64_t kSeqChunk = 1 << 16;
constexpr int64_t kOrdMaxLevels = 1 << 16;
constexpr int64_t kOrdLevelWidth = 16;  // elements per level, on average, below which the loop runs in order
constexpr int kCtlConflict = 1, kCtlRaw = 2, kCtlChanged = 4;

// W^-1 for an affine output whose axes nest: sorted by decreasing |stride|,
// each stride exceeds the span of all narrower axes, so an address has at
// most one coordinate, found by dividing from the widest axis down.  Every
// region, permutation and reversal of a row-major buffer nests.
struct WriteInverse {
  bool ok = false;
  int64_t base = 0;  // the view's lowest address
  int n = 0;         // axes of extent > 1, widest stride first
  int64_t step[PH_MAX_RANK], ext[PH_MAX_RANK], lex[PH_MAX_RANK];
  bool neg[PH_MAX_RANK];
};

WriteInverse write_inverse(const ph_view_desc &W) {
  WriteInverse v;
  if (W.nstage != 0) return v;
  const ph_affine &a = W.view;
  int64_t lex[PH_MAX_RANK];
  int64_t st = 1;
  for (int i = a.rank - 1; i >= 0; --i) {
    lex[i] = st;
    st *= a.shape[i];
  }
  v.base = a.offset;
  int idx[PH_MAX_RANK];
  for (int i = 0; i < a.rank; ++i) {
    if (a.shape[i] <= 1) continue;
    if (a.stride[i] == 0) return v;
    idx[v.n++] = i;
    if (a.stride[i] < 0) v.base += a.stride[i] * (a.shape[i] - 1);
  }
  std::sort(idx, idx + v.n, [&](int x, int y) { return std::llabs(a.stride[x]) > std::llabs(a.stride[y]); });
  for (int k = 0; k < v.n; ++k) {
    const int i = idx[k];
    v.step[k] = std::llabs(a.stride[i]);
    v.ext[k] = a.shape[i];
    v.lex[k] = lex[i];
    v.neg[k] = a.stride[i] < 0;
  }
  for (int k = 0; k < v.n; ++k) {
    int64_t span = 0;
    for (int j = k + 1; j < v.n; ++j) span += v.step[j] * (v.ext[j] - 1);
    if (v.step[k] <= span) return v;
  }
  v.ok = true;
  return v;
}

// Whether two descriptors address the same element at every coordinate.
bool same_map(const ph_view_desc &a, const ph_view_desc &b) {
  if (a.nstage != 0 || b.nstage != 0 || a.view.rank != b.view.rank || a.view.offset != b.view.offset) return false;
  for (int i = 0; i < a.view.rank; ++i) {
    if (a.view.shape[i] != b.view.shape[i]) return false;
    if (a.view.shape[i] > 1 && a.view.stride[i] != b.view.stride[i]) return false;
  }
  return true;
}

// Lowest and highest element index a descriptor can reach.
void desc_span(const ph_view_desc &d, int64_t *lo, int64_t *hi) {
  if (d.nstage > 0) {
    *lo = 0;
    *hi = d.extent - 1;
    return;
  }
  int64_t l = d.view.offset, h = d.view.offset;
  for (int i = 0; i < d.view.rank; ++i) {
    const int64_t span = (d.view.shape[i] - 1) * d.view.stride[i];
    if (span < 0) l += span;
    else h += span;
  }
  *lo = l;
  *hi = h;
}

bool spans_overlap(const ph_view_desc &a, const ph_view_desc &b) {
  int64_t al, ah, bl, bh;
  desc_span(a, &al, &ah);
  desc_span(b, &bl, &bh);
  return !(ah < bl || bh < al);
}

// The coordinates c0.. of lexicographic index `L` over the view's shape.
std::string emit_coords(const ph_affine &v, const std::string &L, const std::string &sfx) {
  std::string s = fmt("      i64 rr%s = %s; (void)rr%s;\n", sfx.c_str(), L.c_str(), sfx.c_str());
  for (int i = v.rank - 1; i >= 0; --i) {
    if (i == 0) {
      s += fmt("      const i64 c0%s = rr%s;\n", sfx.c_str(), sfx.c_str());
    } else {
      const std::string n = lit_i64(v.shape[i]);
      s += fmt("      const i64 c%d%s = rr%s %% %s; rr%s /= %s;\n", i, sfx.c_str(), sfx.c_str(), n.c_str(), sfx.c_str(),
               n.c_str());
    }
  }
  return s;
}

// The element index descriptor `d` gives the coordinates c0..: the affine
// view, then each reshape/tile stage (delinearise, re-linearise).
std::string emit_addr(const ph_view_desc &d, const std::string &name, const std::string &sfx) {
  std::string s = fmt("      i64 %s = %s", name.c_str(), lit_i64(d.view.offset).c_str());
  for (int i = 0; i < d.view.rank; ++i)
    if (d.view.shape[i] > 1 && d.view.stride[i] != 0)
      s += fmt(" + c%d%s * %s", i, sfx.c_str(), lit_i64(d.view.stride[i]).c_str());
  s += ";\n";
  for (int k = 0; k < d.nstage; ++k) {
    const ph_affine &st = d.stage[k];
    s += fmt("      { i64 ix_ = %s; i64 o_ = %s;\n", name.c_str(), lit_i64(st.offset).c_str());
    for (int j = st.rank - 1; j >= 0; --j) {
      const std::string n = lit_i64(st.shape[j]);
      s += fmt("        { const i64 q_ = ix_ / %s; o_ += (ix_ - q_ * %s) * %s; ix_ = q_; }\n", n.c_str(), n.c_str(),
               lit_i64(st.stride[j]).c_str());
    }
    s += fmt("        %s = o_; }\n", name.c_str());
  }
  return s;
}

// ok<sfx>: address `a` is written by some element; lp<sfx>: that element's
// lexicographic index.
std::string emit_inverse(const WriteInverse &w, const std::string &a, const std::string &sfx) {
  const char *x = sfx.c_str();
  std::string s = fmt("      i64 r%s = %s - %s; bool ok%s = r%s >= 0; i64 lp%s = 0;\n", x, a.c_str(),
                      lit_i64(w.base).c_str(), x, x, x);
  for (int k = 0; k < w.n; ++k) {
    const std::string st = lit_i64(w.step[k]);
    const std::string c = w.neg[k] ? fmt("(%s - q_)", lit_i64(w.ext[k] - 1).c_str()) : std::string("q_");
    s += fmt("      { const i64 q_ = r%s / %s; r%s -= q_ * %s; ok%s = ok%s && q_ < %s; lp%s += %s * %s; }\n", x,
             st.c_str(), x, st.c_str(), x, x, lit_i64(w.ext[k]).c_str(), x, c.c_str(), lit_i64(w.lex[k]).c_str());
  }
  s += fmt("      ok%s = ok%s && r%s == 0;\n", x, x, x);
  return s;
}

const char *raw_type(size_t esize) {
  switch (esize) {
    case 1: return "unsigned char";
    case 2: return "unsigned short";
    case 4: return "unsigned";
    default: return "unsigned long long";
  }
}

const char *kOrdPrelude = R"PH(
struct OrdArgs {
  const void *in[8]; void *out; const void *tp; void *tn; const i64 *q0; i64 *q1; int *flags; i64 iter; i64 l0; i64 l1;
  const i64 *ix[8]; const i64 *wx; i64 *inv; i64 *mm; i64 ibase; i64 ispan;
  i64 *dp; const i64 *dq; i64 *cnt; i64 *lv; int *mk;
};
// Adds 1 to c[key] for every active lane, one atomic per distinct key in
// the wavefront; returns the lane's slot: the counter's old value plus the
// lane's rank among the lanes with its key.
__device__ __forceinline__ i64 ph_agg_add(i64 *c, i64 key) {
  const int lane = threadIdx.x & 63;
  unsigned long long pend = __ballot(1);
  i64 pos = 0;
  while (pend) {
    const int lead = __builtin_ctzll(pend);
    const i64 k = __shfl(key, lead);
    const unsigned long long grp = __ballot(key == k) & pend;
    i64 base = 0;
    if (lane == lead) base = (i64)atomicAdd((unsigned long long *)(c + k), (unsigned long long)__popcll(grp));
    base = __shfl(base, lead);
    if (key == k) pos = base + __popcll(grp & ((1ull << lane) - 1));
    pend &= ~grp;
  }
  return pos;
}
__device__ __forceinline__ bool ph_same(bool a, bool b) { return a == b; }
__device__ __forceinline__ bool ph_same(int a, int b) { return a == b; }
__device__ __forceinline__ bool ph_same(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }
__device__ __forceinline__ bool ph_same(double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); }
__device__ __forceinline__ void ph_stp(unsigned char *p, i64 i, bool v) { p[i] = v ? 1 : 0; }
__device__ __forceinline__ void ph_stp(int *p, i64 i, int v) { p[i] = v; }
__device__ __forceinline__ void ph_stp(float *p, i64 i, float v) { p[i] = v; }
__device__ __forceinline__ void ph_stp(double *p, i64 i, double v) { p[i] = v; }
// control bits of a wavefront: one atomic, skipped when its bits are
// already set (they only ever go from 0 to 1; atomics on the one word
// serialise, about 11 ns each, where every wavefront of a pass sets a bit)
__device__ __forceinline__ void ph_ctl(int *w, int ctl) {
  const unsigned long long b1 = __ballot(ctl & 1), b2 = __ballot(ctl & 2), b4 = __ballot(ctl & 4);
  const int v = (b1 ? 1 : 0) | (b2 ? 2 : 0) | (b4 ? 4 : 0);
  if ((threadIdx.x & 63) == 0 && v && (*(volatile int *)w & v) != v) atomicOr(w, v);
}
)PH";

// What the generated kernels need to know: the program (null: a copy of
// input 0 of `esize`-byte elements), the operands, which inputs read the
// output's buffer, and which operands are addressed through index arrays
// (element L of input j at ix[j][L], the output at wx[L]; W^-1 is then a
// table over the addresses the output spans, inv[a - ibase] = the element
// writing a, or -1).
struct OrdSpec {
  const ph_prog *prog = nullptr;
  int n_in = 1;
  int in_dt[PH_MAX_INPUTS] = {0};
  const ph_view_desc *in[PH_MAX_INPUTS] = {nullptr};
  const ph_view_desc *out = nullptr;
  unsigned hazard = 0;
  size_t esize = 0;
  unsigned indexed = 0;
  bool out_indexed = false;
  bool any_index() const { return indexed != 0 || out_indexed; }
};

// The one input of S that reads the output's buffer, or -1 (none, or more).
int single_hazard(const OrdSpec &S) {
  int h = -1;
  for (int j = 0; j < S.n_in; ++j)
    if (S.hazard >> j & 1) {
      if (h >= 0) return -1;
      h = j;
    }
  return h;
}

// A kernel's source with its index arithmetic in 32 bits: a local
// `typedef int i64` and integer literals without their LL suffix (the caller
// has checked that every index, address and extent fits).  Division by a
// constant then takes a 32-bit multiply-high instead of a 64-bit sequence.
std::string narrow_index(const std::string &k) {
  const size_t open = k.find("{\n");
  if (open == std::string::npos) return k;
  std::string r = k.substr(0, open + 2) + "  typedef int i64;\n" + k.substr(open + 2);
  return std::regex_replace(r, std::regex("([0-9])LL"), "$1");
}

// Whether every index and address of S fits in 30 bits (no index arrays).
bool ord_small(const OrdSpec &S) {
  const int64_t lim = int64_t(1) << 30;
  if (S.any_index() || ph::desc_count(S.out) >= lim || S.out->extent >= lim) return false;
  for (int j = 0; j < S.n_in; ++j)
    if (S.in[j]->extent >= lim) return false;
  return true;
}

// part 0: every kernel but those of trees of dependents; part 1: those.
std::string ord_source(const OrdSpec &S, const WriteInverse &w, int part = 0) {
  const int64_t n = ph::desc_count(S.out);
  const std::string N = lit_i64(n);
  const bool copy = S.prog == nullptr;
  const std::string OST = copy ? raw_type(S.esize) : stype(S.out->dtype);
  std::string s = kPrelude;
  s += kOrdPrelude;
  const std::string loop = "    for (i64 L = (i64)blockIdx.x * 256 + threadIdx.x; L < " + N +
                           "; L += (i64)gridDim.x * 256) {\n";
  auto in_st = [&](int j) { return copy ? std::string(raw_type(S.esize)) : std::string(stype(S.in_dt[j])); };
  auto load = [&](int j, const std::string &ptr, const std::string &idx) {
    return copy ? fmt("((const %s *)%s)[%s]", in_st(j).c_str(), ptr.c_str(), idx.c_str())
                : fmt("ph_ld((const %s *)%s, %s)", in_st(j).c_str(), ptr.c_str(), idx.c_str());
  };
  auto vtype = [&](int j) { return copy ? std::string(raw_type(S.esize)) : std::string(ctype(compute_dt(S.in_dt[j]))); };
  // a_j: the address input j reads for element Lv (coordinates c0.. of Lv
  // in scope for a descriptor)
  auto in_addr = [&](int j, const std::string &Lv) {
    if (S.indexed >> j & 1) return fmt("      const i64 a_%d = A.ix[%d][%s];\n", j, j, Lv.c_str());
    return emit_addr(*S.in[j], fmt("a_%d", j), "");
  };
  // ok_j / lp_j: whether a_j is written by some element, and which
  auto inverse = [&](int j) {
    if (S.out_indexed)
      return fmt("      const i64 lp_%d = (a_%d >= A.ibase && a_%d - A.ibase < A.ispan) ? A.inv[a_%d - A.ibase] : -1; "
                 "const bool ok_%d = lp_%d >= 0;\n", j, j, j, j, j, j);
    return emit_inverse(w, fmt("a_%d", j), fmt("_%d", j));
  };
  // w_: the address element Lv writes
  auto out_addr = [&](const std::string &Lv) {
    if (S.out_indexed) return fmt("      const i64 w_ = A.wx[%s];\n", Lv.c_str());
    return emit_addr(*S.out, "w_", "");
  };
  // element program over x0.. (a copy: x0 itself); defines `res`
  auto body = [&](std::string *code) {
    if (copy) {
      *code += "      const auto res = x0;\n";
      return;
    }
    std::vector<std::string> inexpr;
    for (int j = 0; j < S.n_in; ++j) inexpr.push_back(fmt("x%d", j));
    std::string b;
    const std::string r = emit_prog(S.prog, inexpr, &b, "");
    *code += "      const i64 ph_lidx = L; (void)ph_lidx;\n";
    *code += b;
    *code += fmt("      const auto res = %s;\n", r.c_str());
  };

  const bool small = ord_small(S);
  // classify: conflicts and reads after writes
  size_t mark = s.size();
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_classify(OrdArgs A) {\n  int ctl = 0;\n";
  s += loop + emit_coords(S.out->view, "L", "");
  for (int j = 0; j < S.n_in; ++j) {
    if (!(S.hazard >> j & 1)) continue;
    const std::string sx = fmt("_%d", j);
    s += in_addr(j, "L") + inverse(j);
    s += fmt("      if (ok%s && lp%s != L) ctl |= 1; if (ok%s && lp%s < L) ctl |= 2;\n", sx.c_str(), sx.c_str(),
             sx.c_str(), sx.c_str());
  }
  s += "    }\n  ph_ctl(A.flags + 1, ctl);\n}\n";
  if (small) s = s.substr(0, mark) + narrow_index(s.substr(mark));

  // relax: one Jacobi pass into tn, reading tp where an earlier element wrote
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_relax(OrdArgs A) {\n  int fl = 0, ctl = 0;\n";
  s += loop + emit_coords(S.out->view, "L", "");
  for (int j = 0; j < S.n_in; ++j) {
    const std::string sx = fmt("_%d", j);
    s += in_addr(j, "L");
    if (S.hazard >> j & 1) {
      s += inverse(j);
      s += fmt("      const bool dep%s = ok%s && lp%s < L; if (dep%s) ctl |= 2;\n", sx.c_str(), sx.c_str(), sx.c_str(),
               sx.c_str());
      s += fmt("      const %s x%d = (A.iter > 0 && dep%s) ? %s : %s;\n", vtype(j).c_str(), j, sx.c_str(),
               load(j, "A.tp", "lp" + sx).c_str(), load(j, fmt("A.in[%d]", j), "a" + sx).c_str());
    } else {
      s += fmt("      const %s x%d = %s;\n", vtype(j).c_str(), j, load(j, fmt("A.in[%d]", j), "a" + sx).c_str());
    }
  }
  body(&s);
  if (copy) {
    s += fmt("      if (A.iter > 0 && res != ((const %s *)A.tp)[L]) ctl |= 4;\n", OST.c_str());
    s += fmt("      ((%s *)A.tn)[L] = res;\n", OST.c_str());
  } else {
    s += fmt("      if (A.iter > 0 && !ph_same(res, ph_ld((const %s *)A.tp, L))) ctl |= 4;\n", OST.c_str());
    s += fmt("      ph_stp((%s *)A.tn, L, res);\n", OST.c_str());
  }
  s += "    }\n  if (fl) atomicOr(A.flags, fl);\n  ph_ctl(A.flags + 1, ctl);\n}\n";

  // write: out[W(L)] = tp[L]
  mark = s.size();
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_write(OrdArgs A) {\n";
  s += loop + emit_coords(S.out->view, "L", "") + out_addr("L");
  s += fmt("      ((%s *)A.out)[w_] = ((const %s *)A.tp)[L];\n    }\n}\n", OST.c_str(), OST.c_str());
  if (small) s = s.substr(0, mark) + narrow_index(s.substr(mark));

  // seq: the loop itself, in order, over [l0, l1), one thread
  s += "extern \"C\" __global__ void __launch_bounds__(64) ph_ord_seq(OrdArgs A) {\n";
  s += "  if (blockIdx.x != 0 || threadIdx.x != 0) return;\n  int fl = 0;\n";
  s += "    for (i64 L = A.l0; L < A.l1; ++L) {\n" + emit_coords(S.out->view, "L", "");
  for (int j = 0; j < S.n_in; ++j) {
    const std::string sx = fmt("_%d", j);
    s += in_addr(j, "L");
    s += fmt("      const %s x%d = %s;\n", vtype(j).c_str(), j, load(j, fmt("A.in[%d]", j), "a" + sx).c_str());
  }
  body(&s);
  s += out_addr("L");
  if (copy) s += fmt("      ((%s *)A.out)[w_] = res;\n", OST.c_str());
  else s += fmt("      ph_stp((%s *)A.out, w_, res);\n", OST.c_str());
  s += "    }\n  if (fl) atomicOr(A.flags, fl);\n}\n";

  if (copy) {
    // chain links: q1[L] = the earlier element whose write L reads, or -1
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_link(OrdArgs A) {\n";
    s += loop + emit_coords(S.out->view, "L", "") + in_addr(0, "L") + inverse(0);
    s += "      A.q1[L] = (ok_0 && lp_0 < L) ? lp_0 : -1;\n    }\n}\n";
    // one pointer-jumping pass: q1[L] = q0[q0[L]] where both are links
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_jump(OrdArgs A) {\n  int ctl = 0;\n";
    s += loop + "      const i64 q = A.q0[L]; i64 nq = q;\n";
    s += "      if (q >= 0) { const i64 qq = A.q0[q]; if (qq >= 0) { nq = qq; ctl |= 4; } }\n";
    s += "      A.q1[L] = nq;\n    }\n  ph_ctl(A.flags + 1, ctl);\n}\n";
    // gather: tn[L] = the old value at the read address of L's root
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_gather(OrdArgs A) {\n";
    s += loop + "      const i64 q = A.iter ? A.q0[L] : -1; const i64 R = q >= 0 ? q : L;\n";
    s += emit_coords(S.out->view, "R", "") + in_addr(0, "R");
    s += fmt("      ((%s *)A.tn)[L] = ((const %s *)A.in[0])[a_0];\n    }\n}\n", OST.c_str(), OST.c_str());
    const WriteInverse ri = (S.in[0]->nstage == 0 && !S.any_index() && w.ok) ? write_inverse(*S.in[0]) : WriteInverse();
    if (ri.ok) {
      // walkc: a source that reads each address at most once makes the
      // links chains in closed form (roots: no earlier element wrote the
      // address read; successor: the element reading the address written,
      // if later); every element of a chain takes its root's old value
      const size_t mk = s.size();
      s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_walkc(OrdArgs A) {\n" + loop;
      s += emit_coords(S.out->view, "L", "") + in_addr(0, "L") + inverse(0);
      s += "      if (ok_0 && lp_0 < L) continue;\n";
      s += fmt("      const %s v = ((const %s *)A.in[0])[a_0];\n", OST.c_str(), OST.c_str());
      s += "      for (i64 E = L;;) {\n";
      s += fmt("      ((%s *)A.tn)[E] = v;\n", OST.c_str());
      s += emit_coords(S.out->view, "E", "") + out_addr("E") + emit_inverse(ri, "w_", "_s");
      s += "      if (!ok_s || lp_s <= E) break;\n      E = lp_s;\n      }\n    }\n}\n";
      if (small) s = s.substr(0, mk) + narrow_index(s.substr(mk));
    }
  }
  const int hz1 = single_hazard(S);
  std::string t;  // the kernels of trees of dependents (part 1)
  if (S.prog && hz1 >= 0) {
    // Chains of a map with one input on the output's buffer: element L
    // depends on at most the one earlier element that wrote the address
    // it reads (q0[L], or -1).  When no element has two dependents the
    // dependencies form chains, and a thread per chain root runs its chain
    // in order, carrying each result to the next element as the value it
    // reads; the results go to T (tn), indexed by L.
    const std::string sx = fmt("_%d", hz1);
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_mlink(OrdArgs A) {\n";
    s += loop + emit_coords(S.out->view, "L", "") + in_addr(hz1, "L") + inverse(hz1);
    s += fmt("      A.q1[L] = (ok%s && lp%s < L) ? lp%s : -1;\n    }\n}\n", sx.c_str(), sx.c_str(), sx.c_str());
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_unset(OrdArgs A) {\n" + loop;
    s += "      A.q1[L] = -1;\n    }\n}\n";
    // next pointers: q1[q0[L]] = L; a second dependent of one element is a
    // conflict (the dependencies then form a tree)
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_next(OrdArgs A) {\n  int ctl = 0;\n" + loop;
    s += "      const i64 p = A.q0[L];\n      if (p >= 0) {\n";
    s += "        const unsigned long long old = atomicCAS((unsigned long long *)(A.q1 + p), ~0ull, (unsigned long long)L);\n";
    s += "        if (old != ~0ull) ctl |= 1;\n      }\n    }\n  ph_ctl(A.flags + 1, ctl);\n}\n";
    // walk: q0 the links, q1 the next pointers; every element lies on the
    // chain of exactly one root, and next pointers only increase, so each
    // walk ends
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_walk(OrdArgs A) {\n  int fl = 0;\n" + loop;
    // element E's result into T, the input on the buffer reading `from`
    // where E depends on an earlier element (the program reads the element
    // index as L: give it E)
    auto elem = [&](const std::string &dep, const std::string &from) {
      std::string e = emit_coords(S.out->view, "E", "");
      for (int j = 0; j < S.n_in; ++j) {
        e += in_addr(j, "E");
        const std::string ld = load(j, fmt("A.in[%d]", j), fmt("a_%d", j));
        if (j == hz1)
          e += fmt("      const %s x%d = %s ? %s : %s;\n", vtype(j).c_str(), j, dep.c_str(), from.c_str(), ld.c_str());
        else
          e += fmt("      const %s x%d = %s;\n", vtype(j).c_str(), j, ld.c_str());
      }
      std::vector<std::string> inexpr;
      for (int j = 0; j < S.n_in; ++j) inexpr.push_back(fmt("x%d", j));
      std::string b;
      const std::string r = emit_prog(S.prog, inexpr, &b, "");
      e += "      const i64 ph_lidx = E; (void)ph_lidx;\n" + b + fmt("      const auto res = %s;\n", r.c_str());
      return e + fmt("      ph_stp((%s *)A.tn, E, res);\n", OST.c_str());
    };
    s += "      if (A.q0[L] >= 0) continue;\n";
    s += fmt("      %s carry = %s(0);\n      bool from_chain = false;\n", vtype(hz1).c_str(), vtype(hz1).c_str());
    s += "      for (i64 E = L; E >= 0; E = A.q1[E]) {\n";
    s += elem("from_chain", "carry");
    s += fmt("      carry = (%s)res;\n      from_chain = true;\n      }\n    }\n", vtype(hz1).c_str());
    s += "  if (fl) atomicOr(A.flags, fl);\n}\n";
    const WriteInverse ri = (S.in[hz1]->nstage == 0 && !(S.indexed >> hz1 & 1) && !S.out_indexed && w.ok)
                                ? write_inverse(*S.in[hz1])
                                : WriteInverse();
    if (ri.ok) {
      // walkc: the same walk when the input on the buffer reads each
      // address at most once (an injective affine map): the dependencies
      // are then chains, an element is a root when no earlier element
      // wrote the address it reads, and its successor is the element that
      // reads the address it writes (the read map inverted), if that comes
      // later.  No link passes and no pointer loads.
      const size_t mk = s.size();
      s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_walkc(OrdArgs A) {\n  int fl = 0;\n" + loop;
      s += emit_coords(S.out->view, "L", "") + in_addr(hz1, "L") + inverse(hz1);
      s += fmt("      if (ok_%d && lp_%d < L) continue;\n", hz1, hz1);
      s += fmt("      %s carry = %s(0);\n      bool from_chain = false;\n", vtype(hz1).c_str(), vtype(hz1).c_str());
      s += "      for (i64 E = L;;) {\n";
      s += elem("from_chain", "carry");
      s += fmt("      carry = (%s)res;\n      from_chain = true;\n", vtype(hz1).c_str());
      s += out_addr("E") + emit_inverse(ri, "w_", "_s");
      s += "      if (!ok_s || lp_s <= E) break;\n      E = lp_s;\n      }\n    }\n";
      s += "  if (fl) atomicOr(A.flags, fl);\n}\n";
      if (small) s = s.substr(0, mk) + narrow_index(s.substr(mk));
    }

    // Trees whose inner elements (those another element reads) form
    // chains: the walk above runs the chains (q1 then links each inner
    // element to its one inner dependent), then one pass computes every
    // other element from its parent's result in T.
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_mclear(OrdArgs A) {\n" + loop;
    t += "      A.mk[L] = 0;\n    }\n}\n";
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_mark(OrdArgs A) {\n" + loop;
    t += "      const i64 p = A.q0[L]; if (p >= 0) A.mk[p] = 1;\n    }\n}\n";
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_inext(OrdArgs A) {\n  int ctl = 0;\n" + loop;
    t += "      const i64 p = A.q0[L];\n      if (p >= 0 && A.mk[L]) {\n";
    t += "        const unsigned long long old = atomicCAS((unsigned long long *)(A.q1 + p), ~0ull, (unsigned long long)L);\n";
    t += "        if (old != ~0ull) ctl |= 1;\n      }\n    }\n  ph_ctl(A.flags + 1, ctl);\n}\n";
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_leaves(OrdArgs A) {\n  int fl = 0;\n" + loop;
    t += "      if (A.mk[L]) continue;\n      const i64 E = L, pE = A.q0[E];\n";
    t += elem("pE >= 0", load(hz1, "A.tn", "pE"));
    t += "    }\n  if (fl) atomicOr(A.flags, fl);\n}\n";

    // Trees (an element read by several later ones).  Each element's depth
    // by pointer jumping over {ancestor, distance} pairs (dq -> dp); the
    // elements counted and placed by depth (cnt: counts, then cursors); then
    // a pass per level, each element reading its parent's result from T.
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_dinit(OrdArgs A) {\n" + loop;
    t += "      const i64 p = A.q0[L]; A.dp[2 * L] = p; A.dp[2 * L + 1] = p >= 0 ? 1 : 0;\n    }\n}\n";
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_djump(OrdArgs A) {\n  int ctl = 0;\n" + loop;
    t += "      i64 a = A.dq[2 * L], d = A.dq[2 * L + 1];\n";
    t += "      if (a >= 0) { d += A.dq[2 * a + 1]; a = A.dq[2 * a]; if (a >= 0) ctl |= 4; }\n";
    t += "      A.dp[2 * L] = a; A.dp[2 * L + 1] = d;\n    }\n  ph_ctl(A.flags + 1, ctl);\n}\n";
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_dmax(OrdArgs A) {\n  i64 m = 0;\n" + loop;
    t += "      const i64 d = A.dq[2 * L + 1]; m = d > m ? d : m;\n    }\n";
    t += "  for (int o = 32; o > 0; o >>= 1) { const i64 m2 = __shfl_xor(m, o); m = m2 > m ? m2 : m; }\n";
    t += "  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax((unsigned long long *)(A.mm + 2), (unsigned long long)m);\n}\n";
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_lcount(OrdArgs A) {\n" + loop;
    t += "      ph_agg_add(A.cnt, A.dq[2 * L + 1]);\n    }\n}\n";
    // (lv preset to -1: a slot left unplaced, or placed outside lv, raises
    // the conflict bit instead of addressing out of bounds)
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_lplace(OrdArgs A) {\n  int ctl = 0;\n" + loop;
    t += "      const i64 pos = ph_agg_add(A.cnt, A.dq[2 * L + 1]);\n";
    t += "      if (pos >= 0 && pos < " + N + ") A.lv[pos] = L; else ctl |= 1;\n    }\n  ph_ctl(A.flags + 1, ctl);\n}\n";
    // one level: the elements lv[l0 .. l1)
    t += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_level(OrdArgs A) {\n  int fl = 0, ctl = 0;\n";
    t += "    for (i64 ph_li = A.l0 + (i64)blockIdx.x * 256 + threadIdx.x; ph_li < A.l1; ph_li += (i64)gridDim.x * 256) {\n";
    t += "      const i64 E = A.lv[ph_li];\n      if (E < 0 || E >= " + N + ") { ctl |= 1; continue; }\n";
    t += "      const i64 pE = A.q0[E];\n";
    t += elem("pE >= 0", load(hz1, "A.tn", "pE"));
    t += "    }\n  if (fl) atomicOr(A.flags, fl);\n  ph_ctl(A.flags + 1, ctl);\n}\n";
  }
  if (part == 1) return std::string(kPrelude) + kOrdPrelude + t;
  if (S.any_index()) {
    // check: every index inside its buffer, and the span the writes cover
    // (mm preset to {extent, -1} by ph_ord_mm_init)
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_check(OrdArgs A) {\n  int ctl = 0;\n";
    s += "  i64 lo = 0x7fffffffffffffffLL, hi = -1;\n" + loop;
    for (int j = 0; j < S.n_in; ++j)
      if (S.indexed >> j & 1)
        s += fmt("      { const i64 a = A.ix[%d][L]; if (a < 0 || a >= %s) ctl |= 1; }\n", j,
                 lit_i64(S.in[j]->extent).c_str());
    if (S.out_indexed)
      s += fmt("      { const i64 a = A.wx[L]; if (a < 0 || a >= %s) ctl |= 1; lo = a < lo ? a : lo; hi = a > hi ? a : hi; }\n",
               lit_i64(S.out->extent).c_str());
    s += "    }\n  ph_ctl(A.flags + 1, ctl);\n";
    if (S.out_indexed) {
      s += "  for (int o = 32; o > 0; o >>= 1) {\n";
      s += "    const i64 l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);\n";
      s += "    lo = l2 < lo ? l2 : lo; hi = h2 > hi ? h2 : hi;\n  }\n";
      s += "  if ((threadIdx.x & 63) == 0 && hi >= 0) {\n";
      s += "    atomicMin((unsigned long long *)A.mm, (unsigned long long)lo);\n";
      s += "    atomicMax((unsigned long long *)(A.mm + 1), (unsigned long long)hi);\n  }\n";
    }
    s += "}\n";
    s += "extern \"C\" __global__ void ph_ord_mm_init(OrdArgs A) {\n";
    s += "  if (blockIdx.x == 0 && threadIdx.x == 0) { A.mm[0] = " + lit_i64(S.out->extent) + "; A.mm[1] = 0; }\n}\n";
    // direct: one parallel pass, for maps in which no element reads another's output
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_direct(OrdArgs A) {\n  int fl = 0;\n";
    s += loop + emit_coords(S.out->view, "L", "");
    for (int j = 0; j < S.n_in; ++j) {
      s += in_addr(j, "L");
      s += fmt("      const %s x%d = %s;\n", vtype(j).c_str(), j, load(j, fmt("A.in[%d]", j), fmt("a_%d", j)).c_str());
    }
    body(&s);
    s += out_addr("L");
    if (copy) s += fmt("      ((%s *)A.out)[w_] = res;\n", OST.c_str());
    else s += fmt("      ph_stp((%s *)A.out, w_, res);\n", OST.c_str());
    s += "    }\n  if (fl) atomicOr(A.flags, fl);\n}\n";
  }
  if (S.out_indexed) {
    // W^-1 as a table: cleared, filled, then checked (an element named
    // twice leaves one of its writers without its entry)
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_inv_clear(OrdArgs A) {\n";
    s += "  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < A.ispan; i += (i64)gridDim.x * 256) A.inv[i] = -1;\n}\n";
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_inv_fill(OrdArgs A) {\n" + loop;
    s += "      A.inv[A.wx[L] - A.ibase] = L;\n    }\n}\n";
    s += "extern \"C\" __global__ void __launch_bounds__(256) ph_ord_inv_check(OrdArgs A) {\n  int ctl = 0;\n" + loop;
    s += "      if (A.inv[A.wx[L] - A.ibase] != L) ctl |= 1;\n    }\n  ph_ctl(A.flags + 1, ctl);\n}\n";
  }
  return s;
}

struct OrdKernels {
  hipFunction_t classify = nullptr, relax = nullptr, write = nullptr, seq = nullptr;
  hipFunction_t link = nullptr, jump = nullptr, gather = nullptr;
  hipFunction_t check = nullptr, direct = nullptr, mm_init = nullptr, inv_clear = nullptr, inv_fill = nullptr,
                inv_check = nullptr;
  hipFunction_t mlink = nullptr, unset = nullptr, next = nullptr, walk = nullptr, walkc = nullptr;
  hipFunction_t dinit = nullptr, djump = nullptr, dmax = nullptr, lcount = nullptr, lplace = nullptr, level = nullptr;
  hipFunction_t mclear = nullptr, mark = nullptr, inext = nullptr, leaves = nullptr;
};

std::map<std::string, OrdKernels> &ord_cache() {
  static std::map<std::string, OrdKernels> c;
  return c;
}

std::string ord_key(const OrdSpec &S) {
  std::string key = S.prog ? "O" + prog_key(S.prog) : fmt("C%zu", S.esize);
  key_add(&key, S.in_dt, sizeof(int) * (size_t)S.n_in);
  key_add(&key, &S.hazard, sizeof(S.hazard));
  key_add(&key, &S.indexed, sizeof(S.indexed));
  key_add(&key, &S.out_indexed, sizeof(S.out_indexed));
  key_add(&key, S.out, sizeof(ph_view_desc));
  for (int j = 0; j < S.n_in; ++j) key_add(&key, S.in[j], sizeof(ph_view_desc));
  return key;
}

int ord_kernels(const OrdSpec &S, const WriteInverse &w, OrdKernels *K) {
  const std::string key = ord_key(S);
  auto hit = ord_cache().find(key);
  if (hit != ord_cache().end()) {
    *K = hit->second;
    return PH_OK;
  }
  const std::string src = ord_source(S, w);
  int rc = ph::jit_kernel(src, "ph_ord_classify", &K->classify);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_relax", &K->relax);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_write", &K->write);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_seq", &K->seq);
  if (!rc && !S.prog) rc = ph::jit_kernel2(src, "ph_ord_link", &K->link);
  if (!rc && !S.prog) rc = ph::jit_kernel2(src, "ph_ord_jump", &K->jump);
  if (!rc && !S.prog) rc = ph::jit_kernel2(src, "ph_ord_gather", &K->gather);
  if (!rc && S.any_index()) rc = ph::jit_kernel2(src, "ph_ord_check", &K->check);
  if (!rc && S.any_index()) rc = ph::jit_kernel2(src, "ph_ord_direct", &K->direct);
  if (!rc && S.any_index()) rc = ph::jit_kernel2(src, "ph_ord_mm_init", &K->mm_init);
  if (!rc && S.out_indexed) rc = ph::jit_kernel2(src, "ph_ord_inv_clear", &K->inv_clear);
  if (!rc && S.out_indexed) rc = ph::jit_kernel2(src, "ph_ord_inv_fill", &K->inv_fill);
  if (!rc && S.out_indexed) rc = ph::jit_kernel2(src, "ph_ord_inv_check", &K->inv_check);
  const bool chains = S.prog && single_hazard(S) >= 0;
  if (!rc && chains) rc = ph::jit_kernel2(src, "ph_ord_mlink", &K->mlink);
  if (!rc && chains) rc = ph::jit_kernel2(src, "ph_ord_unset", &K->unset);
  if (!rc && chains) rc = ph::jit_kernel2(src, "ph_ord_next", &K->next);
  if (!rc && chains) rc = ph::jit_kernel2(src, "ph_ord_walk", &K->walk);
  if (!rc && src.find("ph_ord_walkc") != std::string::npos) rc = ph::jit_kernel2(src, "ph_ord_walkc", &K->walkc);
  if (rc) return rc;
  ord_cache()[key] = *K;
  return PH_OK;
}

// The kernels of trees of dependents, a module of their own that is
// compiled the first time S meets a tree (the chains' kernels are already in
// *K).
int ord_tree_kernels(const OrdSpec &S, const WriteInverse &w, OrdKernels *K) {
  if (K->leaves) return PH_OK;
  const std::string src = ord_source(S, w, 1);
  int rc = ph::jit_kernel(src, "ph_ord_dinit", &K->dinit);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_djump", &K->djump);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_dmax", &K->dmax);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_lcount", &K->lcount);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_lplace", &K->lplace);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_level", &K->level);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_mclear", &K->mclear);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_mark", &K->mark);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_inext", &K->inext);
  if (!rc) rc = ph::jit_kernel2(src, "ph_ord_leaves", &K->leaves);
  if (rc) {
    K->leaves = nullptr;
    return rc;
  }
  ord_cache()[ord_key(S)] = *K;
  return PH_OK;
}

struct OrdArgsHost {
  const void *in[PH_MAX_INPUTS];
  void *out;
  const void *tp;
  void *tn;
  const int64_t *q0;
  int64_t *q1;
  int *flags;
  int64_t iter, l0, l1;
  const int64_t *ix[PH_MAX_INPUTS];
  const int64_t *wx;
  int64_t *inv;
  int64_t *mm;  // the write index's least and greatest address
  int64_t ibase, ispan;  // the addresses inv covers: [ibase, ibase + ispan)
  int64_t *dp;           // {ancestor, distance} pairs written by a depth pass
  const int64_t *dq;     // the pairs it reads
  int64_t *cnt;          // elements per depth, then the placement cursors
  int64_t *lv;           // the elements ordered by depth
  int *mk;               // 1 for an element that another reads
};

int ord_launch(hipFunction_t fn, OrdArgsHost &A, int64_t n, bool one_thread, const char *what) {
  size_t asz = sizeof(A);
  void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &asz, HIP_LAUNCH_PARAM_END};
  int rc = ph::launch_gate(what);
  if (rc) return rc;
  const int64_t blocks = one_thread ? 1 : std::min<int64_t>((n + 255) / 256, 8192);
  PH_HIP(hipModuleLaunchKernel(fn, (unsigned)(blocks < 1 ? 1 : blocks), 1, 1, one_thread ? 64 : 256, 1, 1, 0,
                               ph::stream(), nullptr, cfg));
  return PH_OK;
}

// Clears the two flag words, runs `fn`, and reads them back.
int ord_pass(hipFunction_t fn, OrdArgsHost &A, int64_t n, int *words, const char *what) {
  int rc = ph::clear_i32(A.flags, 2);
  if (!rc) rc = ord_launch(fn, A, n, false, what);
  if (rc) return rc;
  PH_HIP(hipMemcpyAsync(words, A.flags, 2 * sizeof(int), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  return PH_OK;
}

// The sequential loop, in bounded launches; flags as the loop raises them.
int ord_sequential(const OrdKernels &K, OrdArgsHost &A, int64_t n, int32_t *flags_out) {
  int rc = ph::clear_i32(A.flags, 2);
  if (rc) return rc;
  for (int64_t l0 = 0; l0 < n; l0 += kSeqChunk) {
    A.l0 = l0;
    A.l1 = std::min(n, l0 + kSeqChunk);
    rc = ord_launch(K.seq, A, 1, true, "ordered map (sequential)");
    if (rc) return rc;
  }
  int words[2] = {0, 0};
  PH_HIP(hipMemcpyAsync(words, A.flags, 2 * sizeof(int), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  if (flags_out) *flags_out = words[0];
  return PH_OK;
}

// One parallel pass of an indexed map in which no element reads another's
// output; its flags as raised.
int ord_direct(const OrdKernels &K, OrdArgsHost &A, int64_t n, int32_t *flags_out) {
  int words[2] = {0, 0};
  int rc = ord_pass(K.direct, A, n, words, "indexed map");
  if (rc) return rc;
  if (flags_out) *flags_out = words[0];
  return PH_OK;
}

// A tree of dependents (A.q0: each element's parent, or -1) whose inner
// elements (those another element reads) form chains: the chains walked into
// T (next pointers in A.q1, n slots), then the other elements in one pass.
// *exact false, and nothing evaluated, when an inner element has two inner
// dependents.
int ord_inner_chains(const OrdKernels &K, OrdArgsHost &A, int64_t n, void *T, int *flags, bool *exact) {
  *exact = false;
  void *m;
  int rc = ph::scratch((size_t)n * sizeof(int), &m, 9);
  if (rc) return rc;
  A.mk = (int *)m;
  rc = ord_launch(K.mclear, A, n, false, "ordered map (inner elements)");
  if (!rc) rc = ord_launch(K.mark, A, n, false, "ordered map (inner elements)");
  if (!rc) rc = ord_launch(K.unset, A, n, false, "ordered map (inner elements)");
  int words[2] = {0, 0};
  if (!rc) rc = ord_pass(K.inext, A, n, words, "ordered map (inner elements)");
  if (rc) return rc;
  if (words[1] & kCtlConflict) return PH_OK;
  A.tn = T;
  rc = ord_pass(K.walk, A, n, words, "ordered map (chains)");
  if (rc) return rc;
  const int fw = words[0];
  rc = ord_pass(K.leaves, A, n, words, "ordered map (leaves)");
  if (rc) return rc;
  *flags = fw | words[0];
  *exact = true;
  return PH_OK;
}

// A tree of dependents (A.q0: each element's parent, or -1) evaluated level
// by level into T, the elements ordered by depth in `order` (n slots).
// *exact false, and nothing evaluated, when the tree is deeper than
// kOrdMaxLevels or than n / kOrdLevelWidth (a pass per level would then
// cost more than the loop in order).
int ord_levels(const OrdKernels &K, OrdArgsHost &A, int64_t n, void *T, void *order, int *flags, bool *exact) {
  *exact = false;
  void *p[2];
  int rc = ph::scratch((size_t)n * 2 * sizeof(int64_t), &p[0], 9);
  if (!rc) rc = ph::scratch((size_t)n * 2 * sizeof(int64_t), &p[1], 10);
  if (rc) return rc;
  A.dp = (int64_t *)p[0];
  rc = ord_launch(K.dinit, A, n, false, "ordered map (depths)");
  if (rc) return rc;
  int words[2] = {0, 0};
  int pc = 0;
  bool settled = false;
  for (int round = 0; round < 64 && !settled; ++round) {
    A.dq = (const int64_t *)p[pc];
    A.dp = (int64_t *)p[pc ^ 1];
    rc = ord_pass(K.djump, A, n, words, "ordered map (depths)");
    if (rc) return rc;
    pc ^= 1;
    settled = !(words[1] & kCtlChanged);
  }
  if (!settled) return PH_OK;
  A.dq = (const int64_t *)p[pc];
  rc = ph::clear_i32((int *)(A.mm + 2), 2);
  if (!rc) rc = ord_launch(K.dmax, A, n, false, "ordered map (depths)");
  if (rc) return rc;
  int64_t depth = 0;
  PH_HIP(hipMemcpyAsync(&depth, A.mm + 2, sizeof(depth), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  if (depth > kOrdMaxLevels || depth > n / kOrdLevelWidth) return PH_OK;
  const int64_t nl = depth + 1;
  void *c;
  rc = ph::scratch((size_t)nl * 2 * sizeof(int64_t), &c, 11);
  if (rc) return rc;
  A.cnt = (int64_t *)c;
  rc = ph::clear_i32((int *)c, (int)(nl * 4));
  if (!rc) rc = ord_launch(K.lcount, A, n, false, "ordered map (levels)");
  if (rc) return rc;
  std::vector<int64_t> cnt((size_t)nl), off((size_t)nl);
  PH_HIP(hipMemcpyAsync(cnt.data(), c, (size_t)nl * sizeof(int64_t), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  int64_t total = 0;
  for (int64_t k = 0; k < nl; ++k) {
    off[(size_t)k] = total;
    total += cnt[(size_t)k];
  }
  if (total != n) return ph::set_error(PH_E_HIP, "ordered map: %lld elements placed in levels of %lld",
                                       (long long)total, (long long)n);
  A.cnt = (int64_t *)c + nl;  // cursors, from the offsets
  A.lv = (int64_t *)order;
  PH_HIP(hipMemcpyAsync(A.cnt, off.data(), (size_t)nl * sizeof(int64_t), hipMemcpyHostToDevice, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  A.q1 = (int64_t *)order;
  rc = ord_launch(K.unset, A, n, false, "ordered map (levels)");
  if (!rc) rc = ph::clear_i32(A.flags, 2);
  if (!rc) rc = ord_launch(K.lplace, A, n, false, "ordered map (levels)");
  if (rc) return rc;
  A.tn = T;
  for (int64_t k = 0; k < nl; ++k) {
    A.l0 = off[(size_t)k];
    A.l1 = off[(size_t)k] + cnt[(size_t)k];
    rc = ord_launch(K.level, A, cnt[(size_t)k], false, "ordered map (level)");
    if (rc) return rc;
  }
  PH_HIP(hipMemcpyAsync(words, A.flags, 2 * sizeof(int), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  if (words[1] & kCtlConflict) return ph::set_error(PH_E_HIP, "ordered map: an element left out of its level");
  *flags = words[0];
  *exact = true;
  return PH_OK;
}

// Runs S in the reference's order.  *done false: no element reads another's
// output, and the caller's parallel kernels give the same result.  ix / wx:
// the index arrays of the operands S marks as indexed.
int run_ordered(const OrdSpec &S, const void *const *ins, void *out, int32_t *flags_out, bool *done,
                const int64_t *const *ix = nullptr, const int64_t *wx = nullptr) {
  *done = true;
  if (flags_out) *flags_out = 0;
  const int64_t n = ph::desc_count(S.out);
  const WriteInverse w = S.out_indexed ? WriteInverse() : write_inverse(*S.out);
  OrdKernels K;
  int rc = ord_kernels(S, w, &K);
  if (rc) return rc;
  OrdArgsHost A;
  std::memset(&A, 0, sizeof(A));
  for (int j = 0; j < S.n_in; ++j) {
    A.in[j] = ins[j];
    A.ix[j] = (S.indexed >> j & 1) ? ix[j] : nullptr;
  }
  A.out = out;
  A.wx = S.out_indexed ? wx : nullptr;
  void *fw;
  rc = ph::scratch(64, &fw, 3);
  if (rc) return rc;
  A.flags = (int *)fw;
  A.mm = (int64_t *)fw + 2;  // after the two flag words
  int words[2] = {0, 0};
  if (S.any_index()) {
    rc = ord_launch(K.mm_init, A, 1, true, "indexed map (index bounds)");
    if (!rc) rc = ord_pass(K.check, A, n, words, "indexed map (index bounds)");
    if (rc) return rc;
    if (words[1] & kCtlConflict) return ph::set_error(PH_E_INDEX, "an index array addresses outside its buffer");
    if (S.out_indexed) {
      int64_t mm[2];
      PH_HIP(hipMemcpyAsync(mm, A.mm, sizeof(mm), hipMemcpyDeviceToHost, ph::stream()));
      PH_HIP(hipStreamSynchronize(ph::stream()));
      A.ibase = mm[0];
      A.ispan = mm[1] - mm[0] + 1;
      void *inv;
      rc = ph::scratch((size_t)A.ispan * sizeof(int64_t), &inv, 8);
      if (rc) return rc;
      A.inv = (int64_t *)inv;
      rc = ord_launch(K.inv_clear, A, A.ispan, false, "indexed map (inverse)");
      if (!rc) rc = ord_launch(K.inv_fill, A, n, false, "indexed map (inverse)");
      if (!rc) rc = ord_pass(K.inv_check, A, n, words, "indexed map (inverse)");
      if (rc) return rc;
      if (words[1] & kCtlConflict) return ph::set_error(PH_E_ARGUMENT, "the output index names an element twice");
    }
    if (!S.hazard) return ord_direct(K, A, n, flags_out);
  }
  if (!S.out_indexed && !w.ok) return ord_sequential(K, A, n, flags_out);  // an output that does not nest
  rc = ord_pass(K.classify, A, n, words, "ordered map (classify)");
  if (rc) return rc;
  if (!(words[1] & kCtlConflict)) {
    if (S.any_index()) return ord_direct(K, A, n, flags_out);
    *done = false;
    return PH_OK;
  }
  const size_t esize = S.prog ? ph::dtype_size(S.out->dtype) : S.esize;
  void *t[2];
  rc = ph::scratch((size_t)n * esize, &t[0], 4);
  if (!rc) rc = ph::scratch((size_t)n * esize, &t[1], 5);
  if (rc) return rc;
  const bool raw = (words[1] & kCtlRaw) != 0;
  int cur = 0;  // T holding the result
  if (!S.prog) {
    // a copy: chain links, pointer jumping to the roots, one gather (or,
    // with an injective affine source, the chains walked in closed form)
    A.iter = 0;
    if (raw && K.walkc) {
      A.tn = t[0];
      rc = ord_launch(K.walkc, A, n, false, "ordered copy (chains)");
      if (rc) return rc;
      A.tp = t[0];
      return ord_launch(K.write, A, n, false, "ordered copy (write)");
    }
    if (raw) {
      void *q[2];
      rc = ph::scratch((size_t)n * sizeof(int64_t), &q[0], 6);
      if (!rc) rc = ph::scratch((size_t)n * sizeof(int64_t), &q[1], 7);
      if (rc) return rc;
      A.q1 = (int64_t *)q[0];
      rc = ord_launch(K.link, A, n, false, "ordered copy (links)");
      if (rc) return rc;
      int qc = 0;
      for (int round = 0; round < 64; ++round) {
        A.q0 = (const int64_t *)q[qc];
        A.q1 = (int64_t *)q[qc ^ 1];
        rc = ord_pass(K.jump, A, n, words, "ordered copy (jump)");
        if (rc) return rc;
        qc ^= 1;
        if (!(words[1] & kCtlChanged)) break;
      }
      A.q0 = (const int64_t *)q[qc];
      A.iter = 1;
    }
    A.tn = t[0];
    rc = ord_launch(K.gather, A, n, false, "ordered copy (gather)");
    if (rc) return rc;
  } else {
    if (K.walkc && (words[1] & kCtlRaw)) {
      // chains in closed form: one walk into T, then the write
      A.tn = t[0];
      rc = ord_pass(K.walkc, A, n, words, "ordered map (chains)");
      if (rc) return rc;
      if (flags_out) *flags_out = words[0];
      A.tp = t[0];
      return ord_launch(K.write, A, n, false, "ordered map (write)");
    }
    int flags = 0;
    bool exact = false;
    const int max_iter = K.walk ? kOrdMaxIterWalk : kOrdMaxIter;
    for (int it = 0; it <= max_iter; ++it) {
      A.iter = it;
      A.tp = t[cur];
      A.tn = t[cur ^ 1];
      rc = ord_pass(K.relax, A, n, words, "ordered map (pass)");
      if (rc) return rc;
      cur ^= 1;
      flags = words[0];
      if ((it == 0 && !(words[1] & kCtlRaw)) || (it > 0 && !(words[1] & kCtlChanged))) {
        exact = true;
        break;
      }
    }
    if (!exact && K.walk) {
      // chains deeper than the passes reach: one thread per chain, unless
      // some element has two dependents
      void *q[2];
      rc = ph::scratch((size_t)n * sizeof(int64_t), &q[0], 6);
      if (!rc) rc = ph::scratch((size_t)n * sizeof(int64_t), &q[1], 7);
      if (rc) return rc;
      A.q1 = (int64_t *)q[0];
      rc = ord_launch(K.mlink, A, n, false, "ordered map (links)");
      if (rc) return rc;
      A.q0 = (const int64_t *)q[0];
      A.q1 = (int64_t *)q[1];
      rc = ord_launch(K.unset, A, n, false, "ordered map (next)");
      if (!rc) rc = ord_pass(K.next, A, n, words, "ordered map (next)");
      if (rc) return rc;
      if (!(words[1] & kCtlConflict)) {
        A.tn = t[cur];
        rc = ord_pass(K.walk, A, n, words, "ordered map (chains)");
        if (rc) return rc;
        flags = words[0];
        exact = true;
      } else {
        rc = ord_tree_kernels(S, w, &K);
        if (!rc) rc = ord_inner_chains(K, A, n, t[cur], &flags, &exact);
        if (!rc && !exact) rc = ord_levels(K, A, n, t[cur], q[1], &flags, &exact);
        if (rc) return rc;
      }
    }
    if (!exact) return ord_sequential(K, A, n, flags_out);
    if (flags_out) *flags_out = flags;
  }
  A.tp = t[cur];
  return ord_launch(K.write, A, n, false, "ordered map (write)");
}

// The inputs of a map that read the output's buffer through another
// mapping; an input overlapping the output through another base pointer,
// or reading it as another element type, is refused.
int hazard_inputs(int n_in, const void *const *ins, const ph_view_desc *in_descs, const void *out,
                  const ph_view_desc *out_desc, unsigned *mask) {
  *mask = 0;
  const char *ob = (const char *)out;
  const size_t oes = ph::dtype_size(out_desc->dtype);
  for (int j = 0; j < n_in; ++j) {
    const char *ib = (const char *)ins[j];
    const size_t ies = ph::dtype_size(in_descs[j].dtype);
    if (ib != ob) {
      const bool disjoint = ib + (size_t)in_descs[j].extent * ies <= ob || ob + (size_t)out_desc->extent * oes <= ib;
      if (!disjoint)
        return ph::set_error(PH_E_ARGUMENT, "input %d overlaps the output through another base pointer", j);
      continue;
    }
    if (same_map(in_descs[j], *out_desc)) continue;  // each element reads its own output first
    if (in_descs[j].nstage == 0 && !spans_overlap(in_descs[j], *out_desc)) continue;
    if (in_descs[j].dtype != out_desc->dtype)
      return ph::set_error(PH_E_ARGUMENT, "input %d reads the output's buffer as another element type", j);
    *mask |= 1u << j;
  }
  return PH_OK;
}

}  // namespace

namespace ph {

int ordered_copy(void *dst, const ph_view_desc *dd, const void *src, const ph_view_desc *sd, bool *done) {
  *done = false;
  if (same_map(*dd, *sd)) {  // every element copied onto itself
    *done = true;
    return PH_OK;
  }
  if (dd->nstage == 0 && sd->nstage == 0 && !spans_overlap(*dd, *sd)) return PH_OK;
  OrdSpec S;
  S.n_in = 1;
  S.in_dt[0] = sd->dtype;
  S.in[0] = sd;
  S.out = dd;
  S.hazard = 1;
  S.esize = dtype_size(dd->dtype);
  const void *ins[1] = {src};
  return run_ordered(S, ins, dst, nullptr, done);
}

}  // namespace ph

extern "C" {

long ph_jit_compiles(void) { return ph::jit_compiles(); }

int ph_compile_check(const ph_prog *prog, int32_t n_in, const int32_t *in_dtypes, int64_t rows,
                     int64_t cols, int32_t kind) {
  PH_GUARD();
  if (n_in < 1 || n_in > PH_MAX_INPUTS) return ph::set_error(PH_E_ARGUMENT, "%d inputs", n_in);
  int in_dt[PH_MAX_INPUTS];
  for (int j = 0; j < n_in; ++j) in_dt[j] = in_dtypes[j];
  int rc = validate_prog(prog, n_in, in_dt, kind == 0);
  if (rc) return rc;
  if (rows < 1 || cols < 1) return ph::set_error(PH_E_ARGUMENT, "empty shape");
  if (kind == 5) {
    // the general batch's kernels: inputs are the element and the scalar
    if (n_in != 2 || in_dt[0] != in_dt[1] || (in_dt[0] != PH_I32 && in_dt[0] != PH_F32 && in_dt[0] != PH_F64))
      return ph::set_error(PH_E_ARGUMENT, "kind 5 takes two inputs of one dtype (I32, F32 or F64)");
    return ph::gbatch_compile_check(prog, in_dt[0]);
  }
  if (kind == 6) {
    // the kernels of an indexed map: input 0 reads the output's buffer
    // through an index array, and the output is written through one
    if (in_dt[0] != prog->out_dtype) return ph::set_error(PH_E_ARGUMENT, "kind 6: input 0 must have the output's dtype");
    ph_view_desc d[PH_MAX_INPUTS + 1];
    OrdSpec S;
    S.prog = prog;
    S.n_in = n_in;
    for (int j = 0; j <= n_in; ++j) {
      std::memset(&d[j], 0, sizeof(d[j]));
      d[j].dtype = j < n_in ? in_dt[j] : prog->out_dtype;
      d[j].extent = rows * cols;
      d[j].view.rank = 2;
      d[j].view.shape[0] = rows;
      d[j].view.shape[1] = cols;
      d[j].view.stride[0] = cols;
      d[j].view.stride[1] = 1;
      if (j < n_in) {
        S.in_dt[j] = in_dt[j];
        S.in[j] = &d[j];
      }
    }
    S.out = &d[n_in];
    S.hazard = 1;
    S.esize = ph::dtype_size(prog->out_dtype);
    S.indexed = 1;
    S.out_indexed = true;
    int rc6 = ph::jit_compile_only(ord_source(S, WriteInverse()));
    if (!rc6 && single_hazard(S) >= 0) rc6 = ph::jit_compile_only(ord_source(S, WriteInverse(), 1));
    return rc6;
  }
  if (kind == 3 || kind == 4) {
    // the kernels of a map whose input 0 reads the output's buffer
    // transposed (3), or of a copy reversing a buffer onto itself (4)
    if (kind == 3 && rows != cols) return ph::set_error(PH_E_ARGUMENT, "kind 3 needs a square shape");
    ph_view_desc d[PH_MAX_INPUTS + 1];
    OrdSpec S;
    S.n_in = kind == 3 ? n_in : 1;
    for (int j = 0; j <= S.n_in; ++j) {
      std::memset(&d[j], 0, sizeof(d[j]));
      d[j].dtype = j < S.n_in ? in_dt[j] : (kind == 3 ? prog->out_dtype : in_dt[0]);
      d[j].extent = rows * cols;
      d[j].view.rank = 2;
      d[j].view.shape[0] = rows;
      d[j].view.shape[1] = cols;
      d[j].view.stride[0] = cols;
      d[j].view.stride[1] = 1;
    }
    if (kind == 3) {
      d[0].view.stride[0] = 1;
      d[0].view.stride[1] = cols;
    } else {
      d[0].view.offset = rows * cols - 1;
      d[0].view.stride[0] = -cols;
      d[0].view.stride[1] = -1;
    }
    S.prog = kind == 3 ? prog : nullptr;
    for (int j = 0; j < S.n_in; ++j) {
      S.in_dt[j] = in_dt[j];
      S.in[j] = &d[j];
    }
    S.out = &d[S.n_in];
    S.hazard = 1;
    S.esize = ph::dtype_size(in_dt[0]);
    if (kind == 3 && in_dt[0] != prog->out_dtype)
      return ph::set_error(PH_E_ARGUMENT, "kind 3: input 0 must have the output's dtype");
    const WriteInverse wi = write_inverse(*S.out);
    int rc3 = ph::jit_compile_only(ord_source(S, wi));
    if (!rc3 && S.prog && single_hazard(S) >= 0) rc3 = ph::jit_compile_only(ord_source(S, wi, 1));
    return rc3;
  }
  if (kind == 2) {
    if (prog_uses_index(prog)) return ph::set_error(PH_E_ARGUMENT, "the tiled map has no element index");
    ph_view_desc d[PH_MAX_INPUTS + 1];
    const ph_view_desc *dp[PH_MAX_INPUTS + 1];
    for (int j = 0; j <= n_in; ++j) {
      std::memset(&d[j], 0, sizeof(d[j]));
      d[j].dtype = j < n_in ? in_dt[j] : prog->out_dtype;
      d[j].extent = rows * cols;
      d[j].view.rank = 2;
      d[j].view.shape[0] = rows;
      d[j].view.shape[1] = cols;
      d[j].view.stride[0] = j == 0 ? 1 : cols;    // input 0: the transposed view
      d[j].view.stride[1] = j == 0 ? rows : 1;
      dp[j] = &d[j];
    }
    ph::Layout L;
    ph::collapse(dp, n_in + 1, &L);
    const int ax = tiled_map_axis(L, n_in, 1u);
    if (ax < 0) return ph::set_error(PH_E_ARGUMENT, "the tiled map does not apply to %lld x %lld", (long long)rows,
                                     (long long)cols);
    int64_t nb;
    return ph::jit_compile_only(map_tiled_source(prog, n_in, in_dt, L, prog->out_dtype, ax, 1u, &nb));
  }
  if (kind == 0) {
    ph_view_desc d[PH_MAX_INPUTS + 1];
    const ph_view_desc *dp[PH_MAX_INPUTS + 1];
    for (int j = 0; j <= n_in; ++j) {
      std::memset(&d[j], 0, sizeof(d[j]));
      d[j].dtype = j < n_in ? in_dt[j] : prog->out_dtype;
      d[j].extent = rows * cols;
      d[j].view.rank = 2;
      d[j].view.shape[0] = rows;
      d[j].view.shape[1] = cols;
      d[j].view.stride[0] = cols;
      d[j].view.stride[1] = 1;
      dp[j] = &d[j];
    }
    ph::Layout L;
    ph::collapse(dp, n_in + 1, &L);
    const int ept = map_ept(n_in, in_dt);
    const int64_t c = L.rank > 0 ? L.shape[L.rank - 1] : 1;
    int64_t nb;
    std::string src = map_source(prog, n_in, in_dt, L, prog->out_dtype, ept, L.rank > 1 && c < 256, &nb);
    return ph::jit_compile_only(src);
  }
  View2D V;
  V.rows = rows;
  V.cols = cols;
  for (int j = 0; j < n_in; ++j) {
    V.rs[j] = cols;
    V.cs[j] = 1;
    V.off[j] = 0;
  }
  RedPlan P;
  P.rows = rows;
  P.cols = cols;
  P.vec = 4;
  P.do_cols = P.do_rows = P.do_full = true;
  const int odt = prog->out_dtype;
  if (odt == PH_I32 || odt == PH_BOOL) {
    P.state = "SI"; P.state_bytes = sizeof(HostSI); P.acc_t = "i64";
  } else {
    P.state = "SF"; P.state_bytes = sizeof(HostSF); P.acc_t = "float";
  }
  P.vec = red_vec(std::string(P.state) == "SI", n_in, in_dt);
  for (int j = 0; j < PH_MAX_INPUTS; ++j) P.vec_ok[j] = j < n_in && cols % P.vec == 0;
  P.ntc = (cols + 256 * P.vec - 1) / (256 * P.vec);
  P.ntc_fin = prefolded(P.ntc);
  P.tr = 64;
  P.ntr = (rows + 63) / 64;
  P.ntr_fin = prefolded(P.ntr);
  return ph::jit_compile_only(red_source(prog, n_in, in_dt, V.rs, V.cs, P));
}

int ph_map_indexed(const ph_prog *prog, int32_t n_in, const void *const *ins, const ph_view_desc *in_descs,
                   const int64_t *const *in_index, void *out, const ph_view_desc *out_desc,
                   const int64_t *out_index, int32_t *flags_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (flags_out) *flags_out = 0;
  if (n_in < 1 || n_in > PH_MAX_INPUTS) return ph::set_error(PH_E_ARGUMENT, "%d inputs (1..%d)", n_in, PH_MAX_INPUTS);
  if (!ins || !in_descs || !out_desc) return ph::set_error(PH_E_ARGUMENT, "ph_map_indexed: null argument");
  if (!prog && n_in != 1) return ph::set_error(PH_E_ARGUMENT, "a copy (no program) takes one input");
  // the descriptors the kernels see: an indexed operand keeps its dtype,
  // buffer extent and logical shape
  auto plain = [](ph_view_desc *d) {
    const int r = d->view.rank < 0 ? 0 : std::min(d->view.rank, PH_MAX_RANK);
    d->nstage = 0;
    d->view.offset = 0;
    d->view.pad_ = 0;
    for (int i = 0; i < PH_MAX_RANK; ++i) {
      d->view.stride[i] = 0;
      if (i >= r) d->view.shape[i] = 0;
    }
    std::memset(d->stage, 0, sizeof(d->stage));
  };
  ph_view_desc din[PH_MAX_INPUTS], dout = *out_desc;
  unsigned indexed = 0;
  for (int j = 0; j < n_in; ++j) {
    din[j] = in_descs[j];
    if (in_index && in_index[j]) {
      indexed |= 1u << j;
      plain(&din[j]);
    }
  }
  if (out_index) plain(&dout);
  if (!indexed && !out_index)
    return prog ? ph_map(prog, n_in, ins, in_descs, out, out_desc, flags_out)
                : ph_strided_copy(out, out_desc, ins[0], in_descs, 0);
  int in_dt[PH_MAX_INPUTS];
  int rc;
  if (prog) {
    rc = check_inputs(prog, n_in, ins, din, in_dt, true);
  } else {
    rc = ph::validate_desc(&din[0], "input");
    if (!rc && !ins[0]) rc = ph::set_error(PH_E_ARGUMENT, "input 0: null pointer");
    in_dt[0] = din[0].dtype;
  }
  if (rc) return rc;
  rc = ph::validate_desc(&dout, "output");
  if (rc) return rc;
  if (dout.nstage != 0) return ph::set_error(PH_E_UNSUPPORTED, "map output must be affine or indexed");
  if (prog ? dout.dtype != prog->out_dtype : ph::dtype_size(dout.dtype) != ph::dtype_size(din[0].dtype))
    return ph::set_error(PH_E_ARGUMENT, "output dtype %d does not take the %s", dout.dtype,
                         prog ? "program's result" : "input's elements");
  if (dout.view.rank != din[0].view.rank) return ph::set_error(PH_E_SHAPE, "output rank differs from inputs");
  for (int i = 0; i < dout.view.rank; ++i)
    if (dout.view.shape[i] != din[0].view.shape[i])
      return ph::set_error(PH_E_SHAPE, "output shape differs on axis %d", i);
  const int64_t n = ph::desc_count(&dout);
  if (prog && prog_uses_index(prog) && n > INT32_MAX)
    return ph::set_error(PH_E_OVERFLOW, "element index of %lld elements exceeds Int32", (long long)n);
  if (n == 0) return PH_OK;
  if (!out) return ph::set_error(PH_E_ARGUMENT, "output: null pointer");
  // inputs on the output's buffer (any through an index array, or through
  // a descriptor when the output is indexed) read it in the loop's order
  unsigned hz = 0;
  const size_t oes = ph::dtype_size(dout.dtype);
  for (int j = 0; j < n_in; ++j) {
    const char *ib = (const char *)ins[j], *ob = (const char *)out;
    if (ib != ob) {
      const size_t ies = ph::dtype_size(din[j].dtype);
      if (!(ib + (size_t)din[j].extent * ies <= ob || ob + (size_t)dout.extent * oes <= ib))
        return ph::set_error(PH_E_ARGUMENT, "input %d overlaps the output through another base pointer", j);
      continue;
    }
    if (!(indexed >> j & 1) && !out_index && same_map(din[j], dout)) continue;
    if (prog ? din[j].dtype != dout.dtype : ph::dtype_size(din[j].dtype) != oes)
      return ph::set_error(PH_E_ARGUMENT, "input %d reads the output's buffer as another element type", j);
    hz |= 1u << j;
  }
  OrdSpec S;
  S.prog = prog;
  S.n_in = n_in;
  for (int j = 0; j < n_in; ++j) {
    S.in_dt[j] = in_dt[j];
    S.in[j] = &din[j];
  }
  S.out = &dout;
  S.hazard = hz;
  S.esize = oes;
  S.indexed = indexed;
  S.out_indexed = out_index != nullptr;
  bool done = false;
  return run_ordered(S, ins, out, flags_out, &done, in_index, out_index);
}

int ph_map(const ph_prog *prog, int32_t n_in, const void *const *ins, const ph_view_desc *in_descs,
           void *out, const ph_view_desc *out_desc, int32_t *flags_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (flags_out) *flags_out = 0;
  int in_dt[PH_MAX_INPUTS];
  int rc = check_inputs(prog, n_in, ins, in_descs, in_dt, true);
  if (rc) return rc;
  rc = ph::validate_desc(out_desc, "output");
  if (rc) return rc;
  if (out_desc->nstage != 0) return ph::set_error(PH_E_UNSUPPORTED, "map output must be affine");
  if (out_desc->dtype != prog->out_dtype)
    return ph::set_error(PH_E_ARGUMENT, "output dtype %d, program yields %d", out_desc->dtype, prog->out_dtype);
  if (out_desc->view.rank != in_descs[0].view.rank)
    return ph::set_error(PH_E_SHAPE, "output rank differs from inputs");
  for (int i = 0; i < out_desc->view.rank; ++i)
    if (out_desc->view.shape[i] != in_descs[0].view.shape[i])
      return ph::set_error(PH_E_SHAPE, "output shape differs on axis %d", i);
  if (prog_uses_index(prog) && ph::desc_count(out_desc) > INT32_MAX)
    return ph::set_error(PH_E_OVERFLOW, "element index of %lld elements exceeds Int32",
                         (long long)ph::desc_count(out_desc));
  if (ph::desc_count(out_desc) == 0) return PH_OK;
  // inputs that read the output's buffer through another mapping: the
  // reference's element order decides what they see
  unsigned hz = 0;
  rc = hazard_inputs(n_in, ins, in_descs, out, out_desc, &hz);
  if (rc) return rc;
  if (hz) {
    OrdSpec S;
    S.prog = prog;
    S.n_in = n_in;
    for (int j = 0; j < n_in; ++j) {
      S.in_dt[j] = in_dt[j];
      S.in[j] = &in_descs[j];
    }
    S.out = out_desc;
    S.hazard = hz;
    bool done = false;
    int32_t fl = 0;
    rc = run_ordered(S, ins, out, &fl, &done);
    if (rc) return rc;
    if (done) {
      if (flags_out) *flags_out = fl;
      return PH_OK;
    }
  }
  // inputs with stages are gathered first, and transposed inputs (while
  // the two scratch slots last) are copied through the tiled transpose
  // when the output is contiguous along its innermost axis
  ph_view_desc local[PH_MAX_INPUTS];
  const void *ptrs[PH_MAX_INPUTS];
  int nstaged = 0;
  for (int j = 0; j < n_in; ++j) nstaged += in_descs[j].nstage > 0 ? 1 : 0;
  if (nstaged > 2) return ph::set_error(PH_E_UNSUPPORTED, "at most two reshaped/tiled inputs per map");
  const bool out_inner = inner_contiguous(*out_desc);
  if (out_inner && nstaged == 0 && !prog_uses_index(prog)) {
    // transposed inputs staged through LDS inside the map (at most two)
    unsigned tmask = 0;
    int nt = 0;
    for (int j = 0; j < n_in; ++j)
      if (transposed_layout(in_descs[j])) {
        tmask |= 1u << j;
        ++nt;
      }
    if (nt >= 1 && nt <= 2) {
      const ph_view_desc *td[PH_MAX_INPUTS + 1];
      for (int j = 0; j < n_in; ++j) td[j] = &in_descs[j];
      td[n_in] = out_desc;
      ph::Layout TL;
      ph::collapse(td, n_in + 1, &TL);
      const int ax = tiled_map_axis(TL, n_in, tmask);
      if (ax >= 0) return map_tiled(prog, n_in, ins, in_dt, TL, ax, tmask, out, out_desc, flags_out);
    }
  }
  int slot = 0, spare = 2 - nstaged;
  for (int j = 0; j < n_in; ++j) {
    local[j] = in_descs[j];
    ptrs[j] = ins[j];
    bool copy = in_descs[j].nstage > 0;
    if (!copy && out_inner && spare > 0 && transposed_layout(in_descs[j])) {
      copy = true;
      --spare;
    }
    if (copy) {
      rc = materialize(&in_descs[j], ins[j], slot++, &local[j], &ptrs[j]);
      if (rc) return rc;
    }
  }
  const ph_view_desc *descs[PH_MAX_INPUTS + 1];
  for (int j = 0; j < n_in; ++j) descs[j] = &local[j];
  descs[n_in] = out_desc;
  ph::Layout L;
  ph::collapse(descs, n_in + 1, &L);
  int ept = map_ept(n_in, in_dt);
  if (const char *e = std::getenv("PH_MAP_EPT")) ept = std::min(16, std::max(1, std::atoi(e)));  // tuning override
  const int64_t cols = L.rank > 0 ? L.shape[L.rank - 1] : 1;
  const bool flat = L.rank > 1 && cols < 256;
  int64_t nblocks = 0;
  std::string key = "M" + prog_key(prog);
  key_add(&key, in_dt, sizeof(int) * (size_t)n_in);
  key_add(&key, &L.rank, sizeof(L.rank));
  key_add(&key, L.shape, sizeof(int64_t) * (size_t)L.rank);
  for (int o = 0; o <= n_in; ++o) key_add(&key, L.stride[o], sizeof(int64_t) * (size_t)L.rank);
  const int32_t mk[3] = {out_desc->dtype, ept, flat ? 1 : 0};
  key_add(&key, mk, sizeof(mk));
  hipFunction_t fn;
  auto hit = fused_cache().find(key);
  if (hit != fused_cache().end()) {
    fn = hit->second.k[0];
    nblocks = hit->second.nblocks;
  } else {
    std::string src = map_source(prog, n_in, in_dt, L, out_desc->dtype, ept, flat, &nblocks);
    rc = ph::jit_kernel(src, "ph_map_kernel", &fn);
    if (rc) return rc;
    fused_cache()[key] = FusedKernels{{fn, nullptr, nullptr}, nblocks};
  }
  const bool may_flag = prog_may_flag(prog);
  int *dflags;
  rc = get_flags_buffer(&dflags, may_flag);
  if (rc) return rc;
  MapArgs A;
  std::memset(&A, 0, sizeof(A));
  for (int j = 0; j < n_in; ++j) {
    A.in[j] = ptrs[j];
    A.in_off[j] = L.offset[j];
  }
  A.out = out;
  A.out_off = L.offset[n_in];
  A.flags = dflags;
  size_t asz = sizeof(A);
  void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &asz,
                 HIP_LAUNCH_PARAM_END};
  rc = ph::launch_gate("map");
  if (rc) return rc;
  PH_HIP(hipModuleLaunchKernel(fn, (unsigned)grid_for(nblocks), 1, 1, kBlock, 1, 1, 0, ph::stream(),
                               nullptr, cfg));
  return read_flags(dflags, may_flag, flags_out);
}

// `colex`: the caller's view is the transpose of in_descs, so its
// lexicographic order runs down their columns.
static int reduce2d_all_impl(const ph_prog *prog, int32_t n_in, const void *const *ins,
                             const ph_view_desc *in_descs, int32_t acc, double *full_sum, double *full_max,
                             double *row_sum, double *row_max, double *col_sum, double *col_max,
                             int32_t *flags_out, bool colex) {
  if (flags_out) *flags_out = 0;
  int in_dt[PH_MAX_INPUTS];
  int rc = check_inputs(prog, n_in, ins, in_descs, in_dt);
  if (rc) return rc;
  if (in_descs[0].view.rank != 2) return ph::set_error(PH_E_DIMENSION, "ph_reduce2d_all needs a 2-D view");
  if (ph::desc_count(&in_descs[0]) == 0)
    return ph::set_error(PH_E_ARGUMENT, "reduction over an empty view");
  if (swap_transposed_2d(prog, n_in, in_descs)) {
    ph_view_desc sw[PH_MAX_INPUTS];
    for (int j = 0; j < n_in; ++j) sw[j] = swapped_2d(in_descs[j]);
    return reduce2d_all_impl(prog, n_in, ins, sw, acc, full_sum, full_max, col_sum, col_max, row_sum, row_max,
                             flags_out, !colex);
  }
  View2D V;
  const void *ptrs[PH_MAX_INPUTS];
  ph_view_desc local[PH_MAX_INPUTS];
  V.rows = in_descs[0].view.shape[0];
  V.cols = in_descs[0].view.shape[1];
  for (int j = 0; j < n_in; ++j) {
    local[j] = in_descs[j];
    ptrs[j] = ins[j];
    if (in_descs[j].nstage > 0 || (j <= 1 && transposed_layout(in_descs[j]))) {
      if (j > 1) return ph::set_error(PH_E_UNSUPPORTED, "at most two reshaped/tiled inputs");
      rc = materialize(&in_descs[j], ins[j], j, &local[j], &ptrs[j]);
      if (rc) return rc;
    }
    V.rs[j] = local[j].view.stride[0];
    V.cs[j] = local[j].view.stride[1];
    V.off[j] = local[j].view.offset;
  }
  return run_reduce2d(prog, n_in, ptrs, in_dt, V, acc, full_sum, full_max, row_sum, row_max, col_sum,
                      col_max, flags_out, colex);
}

int ph_reduce2d_all(const ph_prog *prog, int32_t n_in, const void *const *ins,
                    const ph_view_desc *in_descs, int32_t acc, double *full_sum, double *full_max,
                    double *row_sum, double *row_max, double *col_sum, double *col_max,
                    int32_t *flags_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  return reduce2d_all_impl(prog, n_in, ins, in_descs, acc, full_sum, full_max, row_sum, row_max, col_sum, col_max,
                           flags_out, false);
}

static int reduce_impl(const ph_prog *prog, int32_t n_in, const void *const *ins, const ph_view_desc *in_descs,
                       int32_t keep_axis, int32_t ops, int32_t acc, double *out_sum, double *out_max,
                       int32_t *flags_out, bool colex) {
  if (flags_out) *flags_out = 0;
  int in_dt[PH_MAX_INPUTS];
  int rc = check_inputs(prog, n_in, ins, in_descs, in_dt);
  if (rc) return rc;
  const int rank = in_descs[0].view.rank;
  if (rank < 1) return ph::set_error(PH_E_DIMENSION, "reduction of a rank-0 view");
  if (keep_axis < -1 || keep_axis >= rank)
    return ph::set_error(PH_E_INDEX, "axis %d outside a %d-dimensional view", keep_axis, rank);
  double *sum = (ops & PH_RED_SUM) ? out_sum : nullptr;
  double *mx = (ops & PH_RED_MAX) ? out_max : nullptr;
  if (ph::desc_count(&in_descs[0]) == 0) {
    if (keep_axis >= 0 && in_descs[0].view.shape[keep_axis] == 0) return PH_OK;  // no results
    return ph::set_error(PH_E_ARGUMENT, "reduction over an empty view");
  }
  if (rank == 2 && swap_transposed_2d(prog, n_in, in_descs)) {
    ph_view_desc sw[PH_MAX_INPUTS];
    for (int j = 0; j < n_in; ++j) sw[j] = swapped_2d(in_descs[j]);
    return reduce_impl(prog, n_in, ins, sw, keep_axis < 0 ? keep_axis : 1 - keep_axis, ops, acc, out_sum, out_max,
                       flags_out, !colex);
  }
  // Arrange as 2-D: rows x cols in lexicographic order of the reduced data.
  //   keep -1         : [prod(shape[:-1]), shape[-1]]      -> full
  //   keep 0          : [shape[0], prod(shape[1:])]         -> per row
  //   keep rank-1     : [prod(shape[:-1]), shape[-1]]       -> per column
  //   keep k (middle) : axis k moved to the front (gathered) -> per row
  ph_view_desc local[PH_MAX_INPUTS];
  const void *ptrs[PH_MAX_INPUTS];
  for (int j = 0; j < n_in; ++j) {
    local[j] = in_descs[j];
    ptrs[j] = ins[j];
  }
  bool middle = keep_axis > 0 && keep_axis < rank - 1;
  bool col_mode = keep_axis == rank - 1 && rank > 1;
  bool row_mode = keep_axis == 0 || middle || (keep_axis == 0 && rank == 1);
  if (keep_axis == 0 && rank == 1) {
    // each slice is one element: a map of the program
    row_mode = true;
  }
  for (int j = 0; j < n_in; ++j) {
    ph_view_desc d = local[j];
    bool need_copy = d.nstage > 0;
    if (middle) {
      // permute axis k to the front (a pure view change; for a staged view
      // too, whose stages act on the linear index after the view's map)
      ph_view_desc p = d;
      p.view.shape[0] = d.view.shape[keep_axis];
      p.view.stride[0] = d.view.stride[keep_axis];
      int o = 1;
      for (int i = 0; i < rank; ++i)
        if (i != keep_axis) {
          p.view.shape[o] = d.view.shape[i];
          p.view.stride[o] = d.view.stride[i];
          ++o;
        }
      d = p;
    }
    if (!need_copy) {
      // can the needed 2-D split be expressed by strides?
      int split = (keep_axis == -1 || col_mode) ? rank - 1 : 1;  // dims [0,split) -> rows
      for (int i = 0; i + 1 < split; ++i)
        if (d.view.shape[i + 1] != 1 && d.view.stride[i] != d.view.stride[i + 1] * d.view.shape[i + 1])
          need_copy = true;
      for (int i = split; i + 1 < rank; ++i)
        if (d.view.shape[i + 1] != 1 && d.view.stride[i] != d.view.stride[i + 1] * d.view.shape[i + 1])
          need_copy = true;
    }
    if (!need_copy && j <= 1 && transposed_layout(d)) need_copy = true;
    if (need_copy) {
      if (j > 1) return ph::set_error(PH_E_UNSUPPORTED, "at most two non-collapsible inputs");
      rc = materialize(&d, ptrs[j], j, &local[j], &ptrs[j]);
      if (rc) return rc;
    } else {
      local[j] = d;
    }
  }
  View2D V;
  const int split = (keep_axis == -1 || col_mode) ? rank - 1 : 1;
  V.rows = 1;
  V.cols = 1;
  for (int i = 0; i < split; ++i) V.rows *= local[0].view.shape[i];
  for (int i = split; i < rank; ++i) V.cols *= local[0].view.shape[i];
  for (int j = 0; j < n_in; ++j) {
    const ph_affine &a = local[j].view;
    // innermost non-unit stride of each group (any stride if the group has one element)
    int64_t rs = 0, cs = 0;
    for (int i = split - 1; i >= 0; --i)
      if (a.shape[i] != 1) { rs = a.stride[i]; break; }
    for (int i = rank - 1; i >= split; --i)
      if (a.shape[i] != 1) { cs = a.stride[i]; break; }
    if (rank == 1 && split == 0) { V.rows = 1; rs = 0; }
    V.rs[j] = rs;
    V.cs[j] = cs;
    V.off[j] = a.offset;
  }
  if (rank == 1 && keep_axis == -1) {
    // single row
    V.rows = 1;
    V.cols = local[0].view.shape[0];
    for (int j = 0; j < n_in; ++j) {
      V.rs[j] = 0;
      V.cs[j] = local[j].view.stride[0];
    }
  }
  if (rank == 1 && keep_axis == 0) {
    V.rows = local[0].view.shape[0];
    V.cols = 1;
    for (int j = 0; j < n_in; ++j) {
      V.rs[j] = local[j].view.stride[0];
      V.cs[j] = 0;
    }
  }
  if (keep_axis == -1)
    return run_reduce2d(prog, n_in, ptrs, in_dt, V, acc, sum, mx, nullptr, nullptr, nullptr, nullptr,
                        flags_out, colex);
  if (col_mode)
    return run_reduce2d(prog, n_in, ptrs, in_dt, V, acc, nullptr, nullptr, nullptr, nullptr, sum, mx,
                        flags_out, colex);
  (void)row_mode;
  return run_reduce2d(prog, n_in, ptrs, in_dt, V, acc, nullptr, nullptr, sum, mx, nullptr, nullptr,
                      flags_out, colex);
}

int ph_reduce(const ph_prog *prog, int32_t n_in, const void *const *ins, const ph_view_desc *in_descs,
              int32_t keep_axis, int32_t ops, int32_t acc, double *out_sum, double *out_max,
              int32_t *flags_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  return reduce_impl(prog, n_in, ins, in_descs, keep_axis, ops, acc, out_sum, out_max, flags_out, false);
}

}  // extern "C"
