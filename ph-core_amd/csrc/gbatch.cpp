// General batches of small arrays (ph_gbatch_*).  Each record names one
// array of a packed buffer and the work on it: an element program whose
// results go to y, another whose values are summed and maximised, and a
// strided region copied into z.  One persistent kernel does all of it in one
// pass over x; hipRTC generates it for the batch's programs, with one
// streaming loop per distinct (map program, reduction program) pair among
// the records (chosen once per work item, outside the element loops).
// Whether an item sums, maximises or copies a region is a per-item value,
// the same for the whole wavefront, so work it does not ask for is skipped
// by scalar branches.
//
// A wavefront runs one work item (an array, or a chunk of a large one) at a
// time: it reads its elements once with 16-byte loads, kSub elements per
// step, evaluates the programs per element, and keeps the step in LDS when
// the record copies a region.  The region is normalised to positive steps
// (a negative step becomes a positive one with that axis of z reversed); its
// source elements then increase with the normalised index, so the elements
// sourced from one step form one contiguous range of normalised indices,
// gathered from LDS and scattered into z.  Sums are FP64 per lane, folded
// across the wavefront by a fixed butterfly and across an array's items in
// item order by a second kernel, so they do not depend on which wavefront
// ran which item.  Maxima carry (value, first index): the larger value wins,
// equal values keep the lower index, and a NaN counts only at index 0, where
// it sticks; that state combines in any order and gives exactly the result
// of Enumerable#max's `>` loop (src/multi_indexable.cr:18).  Int32 values
// are summed exactly, with Crystal's checked `+`: a step's elements are taken
// in lexicographic order (vector m, then lane, then element), so a prefix scan
// across the wavefront per vector gives every running total; an item keeps
// its total and its lowest and highest running total, and the fold kernel
// chains the items in order and raises the overflow flag when a running
// total leaves Int32, as the sequential loop would.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "gen.hpp"
#include "jit.hpp"

namespace {

using namespace ph::gen;

constexpr int kMaxProgs = 16;
constexpr int kMaxCombos = 64;

// One work item on the device; the generated source declares the same
// layout (kGRec).
struct GRec {
  int64_t off, zoff;            // the array's element offsets in x/y and z
  uint32_t lo, hi;              // element range [lo, hi) of the array
  uint32_t zlo, zhi;            // normalised region elements sourced from [lo, hi)
  uint32_t shp[3];              // shape, padded on the left with 1
  uint32_t f[3], s[3], cnt[3];  // normalised region: first, step > 0, count
  uint32_t m_c, l_c, m_b, l_b;  // magic divisors for shp[2], shp[1]
  uint32_t mz_c, lz_c, mz_b, lz_b;  // magic divisors for cnt[2], cnt[1]
  uint32_t rev;                 // bit 2 - a: axis a of z runs backwards
  int32_t combo;                // which streaming loop runs the item
  int32_t slot;                 // partial-result slot (array-major item index)
  int32_t ops;                  // PH_RED_SUM | PH_RED_MAX
  double scalar;
};
static_assert(sizeof(GRec) == 136, "GRec layout");

const char *kGRec = R"PH(
typedef float gf4 __attribute__((ext_vector_type(4)));
typedef int gi4 __attribute__((ext_vector_type(4)));
typedef double gd2 __attribute__((ext_vector_type(2)));
struct GRec {
  i64 off, zoff;
  unsigned lo, hi, zlo, zhi;
  unsigned shp[3], f[3], s[3], cnt[3];
  unsigned m_c, l_c, m_b, l_b, mz_c, lz_c, mz_b, lz_b;
  unsigned rev;
  int combo, slot, ops;
  double scalar;
};
struct GArgs {
  const GRec *recs;
  const void *x; void *y; void *z;
  double *psum; double *pmv; unsigned *pmi;
  const int *first; const int *rops;
  double *sums; double *maxes;
  int *counter;
  long long *pint;
  int nitems, dynamic, n, pad_;
};
__device__ __forceinline__ unsigned fdiv(unsigned n, unsigned m, unsigned l) {
  return (unsigned)(((unsigned long long)__umulhi(n, m) + n) >> l);
}
// region elements along one axis whose index f + a s lies below i, and
// whether one lies at i
__device__ __forceinline__ unsigned g_below(unsigned i, unsigned f, unsigned s, unsigned c) {
  if (i <= f) return 0u;
  const unsigned q = (i - f + s - 1u) / s;
  return q < c ? q : c;
}
__device__ __forceinline__ bool g_hit(unsigned i, unsigned f, unsigned s, unsigned c) {
  if (i < f) return false;
  const unsigned d = i - f;
  return d % s == 0u && d / s < c;
}
// normalised region elements whose source index is below e (e < the size)
__device__ unsigned g_slice_before(const GRec &R, unsigned e) {
  const unsigned t = fdiv(e, R.m_c, R.l_c);
  const unsigned k = e - t * R.shp[2];
  const unsigned i = fdiv(t, R.m_b, R.l_b);
  const unsigned j = t - i * R.shp[1];
  unsigned n = g_below(i, R.f[0], R.s[0], R.cnt[0]) * R.cnt[1] * R.cnt[2];
  if (g_hit(i, R.f[0], R.s[0], R.cnt[0])) {
    n += g_below(j, R.f[1], R.s[1], R.cnt[1]) * R.cnt[2];
    if (g_hit(j, R.f[1], R.s[1], R.cnt[1])) n += g_below(k, R.f[2], R.s[2], R.cnt[2]);
  }
  return n;
}
// (value, first index) maximum: see gbatch.cpp
__device__ __forceinline__ void g_max_comb(double &v, unsigned &i, double v2, unsigned i2) {
  if (v != v) return;
  if (v2 != v2 || v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}
)PH";

struct GArgsHost {
  const void *recs;
  const void *x;
  void *y;
  void *z;
  double *psum;
  double *pmv;
  unsigned *pmi;
  const int *first;
  const int *rops;
  double *sums;
  double *maxes;
  int *counter;
  long long *pint;
  int nitems, dynamic, n, pad_;
};

struct Combo {  // a (map program, reduction program) pair, -1 for none
  int map, red;
};

struct GArray {
  int64_t off, zoff, n, zn;
  int rank;
  int64_t shp[3];
  int64_t f[3], s[3], cnt[3];  // normalised region (s > 0)
  uint32_t rev;
  int combo;  // -1: nothing to do
  const ph_batch_rec *rec;
};

struct GBatch {
  int xdt = PH_F32, ydt = PH_F32;
  int n = 0;
  std::vector<ph_prog> progs;
  std::vector<Combo> combos;
  std::vector<GArray> arrays;
  std::vector<ph_batch_rec> recs;
  int64_t total = 0, ztotal = 0;
  bool may_flag = false, any_red = false, any_slice = false;
  std::string src;
  hipFunction_t k_main = nullptr, k_final = nullptr, k_reset = nullptr;
  GRec *d_recs = nullptr;      // work items, largest first (the batched launch)
  GRec *d_recs_am = nullptr;   // the same in array order (one launch per array)
  std::vector<int32_t> first;  // first item of each array in array order (n + 1)
  int32_t *d_first = nullptr, *d_rops = nullptr, *d_counter = nullptr;
  double *d_psum = nullptr, *d_pmv = nullptr;
  unsigned *d_pmi = nullptr;
  long long *d_pint = nullptr;  // Int32 sums: per item total, lowest and highest running total
  bool any_isum = false;
  void *d_scalars = nullptr;  // per record, x dtype (the unbatched form's scalar operand)
  int64_t nitems = 0, items_cap = 0;
  int64_t chunk = 16384;
  int grid = 0, dynamic = 0;
  // the launch arguments and the graph captured with them
  GArgsHost args;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void magic_u32(uint32_t d, uint32_t *m, uint32_t *l) {
  if (d < 1) d = 1;
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  *l = s;
  *m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
}

// Host twin of g_slice_before for e in [0, n].
int64_t slice_before(const GArray &A, int64_t e) {
  if (e >= A.n) return A.zn;
  const int64_t k = e % A.shp[2], t = e / A.shp[2], j = t % A.shp[1], i = t / A.shp[1];
  auto below = [](int64_t x, int64_t f, int64_t s, int64_t c) -> int64_t {
    if (x <= f) return 0;
    return std::min<int64_t>(c, (x - f + s - 1) / s);
  };
  auto hit = [](int64_t x, int64_t f, int64_t s, int64_t c) {
    return x >= f && (x - f) % s == 0 && (x - f) / s < c;
  };
  int64_t cnt = below(i, A.f[0], A.s[0], A.cnt[0]) * A.cnt[1] * A.cnt[2];
  if (hit(i, A.f[0], A.s[0], A.cnt[0])) {
    cnt += below(j, A.f[1], A.s[1], A.cnt[1]) * A.cnt[2];
    if (hit(j, A.f[1], A.s[1], A.cnt[1])) cnt += below(k, A.f[2], A.s[2], A.cnt[2]);
  }
  return cnt;
}

// The 16-byte vector of a 4- or 8-byte dtype (the only x dtypes, and the y
// dtypes stored as vectors).
const char *vec_type(int dt) {
  if (dt == PH_F32) return "gf4";
  if (dt == PH_I32) return "gi4";
  if (dt == PH_F64) return "gd2";
  return "void";
}

// The streaming loop of one combination.
std::string item_source(const GBatch &B, int c, int vec, int msteps) {
  const Combo &C = B.combos[(size_t)c];
  const int sub = 64 * vec * msteps;
  const bool slice = B.any_slice;
  const char *XS = stype(B.xdt);
  const char *XC = ctype(compute_dt(B.xdt));
  const char *YS = stype(B.ydt);
  const char *XV = vec_type(B.xdt);
  const bool ysame = C.map >= 0 && B.ydt != PH_BOOL && ph::dtype_size(B.ydt) == ph::dtype_size(B.xdt);
  const bool red = C.red >= 0;
  const bool ired = red && B.progs[(size_t)C.red].out_dtype == PH_I32;
  // DS / DM: the item sums / maximises (R.ops, dispatched once per item)
  std::string s = fmt("template <bool DS, bool DM>\n__device__ __forceinline__ void item_%d(const GArgs &A, const GRec &R, "
                      "const int lane, %s *sx) {\n", c, XS);
  s += fmt("  const %s *__restrict__ xa = (const %s *)A.x + R.off;\n", XS, XS);
  if (C.map >= 0) s += fmt("  %s *__restrict__ ya = (%s *)A.y + R.off;\n", YS, YS);
  if (slice)
    s += fmt("  %s *__restrict__ za = (%s *)A.z + R.zoff;\n  unsigned zlo = R.zlo;\n  const bool sl = R.zlo < R.zhi;\n",
             XS, XS);
  s += fmt("  const %s sc = (%s)R.scalar;\n  int fl = 0;\n", XC, XC);
  if (red && !ired)
    s += "  double gs = 0.0;\n  double mv = -__builtin_inf(); unsigned mi = 0xffffffffu;\n";
  if (ired)
    s += "  long long it = 0, ilo = 0, ihi = 0;  // total so far; lowest / highest running total\n"
         "  int imx = -2147483647 - 1;\n";
  s += fmt("  for (unsigned base = R.lo; base < R.hi; base += %du) {\n", sub);
  s += fmt("    const unsigned end = min(R.hi, base + %du);\n", sub);
  // lanes past the end load the last whole vector inside the array (its
  // values are discarded), so every load is in bounds and none is skipped
  s += fmt("    const unsigned lastv = (end - 1u) & ~%du;\n", vec - 1);
  s += fmt("    %s v[%d];\n#pragma unroll\n    for (int m = 0; m < %d; ++m) {\n", XV, msteps, msteps);
  s += fmt("      const unsigned e = base + %du * (unsigned)lane + %du * (unsigned)m;\n", vec, 64 * vec);
  s += fmt("      v[m] = *(const %s *)(xa + (e < end ? e : lastv));\n    }\n", XV);
  if (slice)
    s += fmt("    if (sl) {\n#pragma unroll\n      for (int m = 0; m < %d; ++m) ((%s *)sx)[lane + 64 * m] = v[m];\n    }\n",
             msteps, XV);
  s += fmt("#pragma unroll\n    for (int m = 0; m < %d; ++m) {\n", msteps);
  s += fmt("      const unsigned e0 = base + %du * (unsigned)lane + %du * (unsigned)m;\n", vec, 64 * vec);
  s += "      const " + std::string(XS) + " xs_[" + std::to_string(vec) + "] = {";
  for (int j = 0; j < vec; ++j) s += fmt("%sv[m].%c", j ? ", " : "", "xyzw"[j]);
  s += "};\n";
  if (ysame) s += fmt("      %s yo[%d];\n", YS, vec);
  if (ired) s += fmt("      long long iv[%d];\n", vec);
  s += fmt("#pragma unroll\n      for (int j = 0; j < %d; ++j) {\n", vec);
  s += "        const unsigned e = e0 + j;\n        const bool ok = e < end;\n";
  s += fmt("        const %s xe = (%s)xs_[j];\n", XC, XC);
  if (C.map >= 0) {
    std::string body;
    const std::string r = emit_prog(&B.progs[(size_t)C.map], {"xe", "sc"}, &body, "_m");
    // a lane past the item's end computes on stale values: its flags are dropped
    s += "        const int flm_ = fl;\n" + body + "        if (!ok) fl = flm_;\n";
    if (ysame)
      s += fmt("        yo[j] = (%s)(%s);\n", YS, r.c_str());
    else
      s += fmt("        if (ok) ph_st(ya, (i64)e, (%s)(%s));\n", ctype(B.ydt), r.c_str());
  }
  if (C.red >= 0) {
    std::string body;
    const std::string r = emit_prog(&B.progs[(size_t)C.red], {"xe", "sc"}, &body, "_r");
    s += "        const int flr_ = fl;\n" + body + "        if (!ok) fl = flr_;\n";
    if (ired) {
      s += fmt("        iv[j] = ok ? (long long)(%s) : 0LL;\n", r.c_str());
      s += fmt("        if (DM && ok) imx = imx > (%s) ? imx : (%s);\n", r.c_str(), r.c_str());
    } else {
      s += fmt("        const double rv = (double)(%s);\n", r.c_str());
      s += "        if (DS) gs += ok ? rv : 0.0;\n";
      s += "        if (DM && ok) {\n          if (rv != rv) { if (e == 0u) { mv = rv; mi = 0u; } }\n"
           "          else if (rv > mv || (rv == mv && e < mi)) { mv = rv; mi = e; }\n        }\n";
    }
  }
  s += "      }\n";
  if (ired) {
    // running totals of this vector's elements: the lane's own prefix sums
    // plus the wavefront's exclusive prefix of lane totals plus the item's
    // total so far (elements past the end add 0)
    s += fmt("      if (DS) {\n        long long pf = 0, pmin = 0, pmax = 0;\n#pragma unroll\n        for (int j = 0; j < %d; ++j) {\n"
             "          pf += iv[j];\n          pmin = pf < pmin ? pf : pmin;\n          pmax = pf > pmax ? pf : pmax;\n        }\n",
             vec);
    s += "        long long inc = pf;\n"
         "        for (int d = 1; d < 64; d <<= 1) {\n          const long long t = __shfl_up(inc, d, 64);\n"
         "          if (lane >= d) inc += t;\n        }\n"
         "        const long long ex = it + inc - pf;\n"
         "        ilo = ex + pmin < ilo ? ex + pmin : ilo;\n        ihi = ex + pmax > ihi ? ex + pmax : ihi;\n"
         "        it += __shfl(inc, 63, 64);\n      }\n";
  }
  if (ysame) {
    const char *YV = vec_type(B.ydt);
    s += fmt("      if (e0 + %du <= end) {\n        %s yw;\n", vec, YV);
    for (int j = 0; j < vec; ++j) s += fmt("        yw.%c = yo[%d];\n", "xyzw"[j], j);
    s += fmt("        __builtin_nontemporal_store(yw, (%s *)(ya + e0));\n      } else {\n", YV);
    s += fmt("        for (int j = 0; j < %d; ++j) if (e0 + j < end) ya[e0 + j] = yo[j];\n      }\n", vec);
  }
  s += "    }\n";
  if (slice) {
    s += "    if (sl) {\n    __builtin_amdgcn_wave_barrier();\n";
    s += "    const unsigned zhi = end == R.hi ? R.zhi : g_slice_before(R, end);\n";
    s += "    for (unsigned z0 = zlo; z0 < zhi; z0 += 256u) {\n";
    s += fmt("      %s g[4]; unsigned dz[4];\n#pragma unroll\n      for (int u = 0; u < 4; ++u) {\n", XS);
    s += "        const unsigned zi = z0 + 64u * u + (unsigned)lane;\n        dz[u] = 0xffffffffu;\n";
    s += "        if (zi < zhi) {\n";
    s += "          const unsigned t = fdiv(zi, R.mz_c, R.lz_c);\n          const unsigned c = zi - t * R.cnt[2];\n";
    s += "          const unsigned a = fdiv(t, R.mz_b, R.lz_b);\n          const unsigned b = t - a * R.cnt[1];\n";
    s += "          const unsigned src = ((R.f[0] + a * R.s[0]) * R.shp[1] + R.f[1] + b * R.s[1]) * R.shp[2] + R.f[2] + c * R.s[2];\n";
    s += "          g[u] = sx[src - base];\n";
    s += "          const unsigned aa = (R.rev & 4u) ? R.cnt[0] - 1u - a : a;\n";
    s += "          const unsigned bb = (R.rev & 2u) ? R.cnt[1] - 1u - b : b;\n";
    s += "          const unsigned cc = (R.rev & 1u) ? R.cnt[2] - 1u - c : c;\n";
    s += "          dz[u] = (aa * R.cnt[1] + bb) * R.cnt[2] + cc;\n        }\n      }\n";
    s += "#pragma unroll\n      for (int u = 0; u < 4; ++u) if (dz[u] != 0xffffffffu) za[dz[u]] = g[u];\n    }\n";
    s += "    zlo = zhi;\n    __builtin_amdgcn_wave_barrier();  // the next step overwrites the LDS copy\n    }\n";
  }
  s += "  }\n";
  if (ired) {
    s += "  if (DS) {\n    for (int o = 32; o > 0; o >>= 1) {\n"
         "      const long long a2 = __shfl_xor(ilo, o, 64), b2 = __shfl_xor(ihi, o, 64);\n"
         "      ilo = a2 < ilo ? a2 : ilo;\n      ihi = b2 > ihi ? b2 : ihi;\n    }\n"
         "    if (lane == 0) { A.pint[3 * R.slot] = it; A.pint[3 * R.slot + 1] = ilo; A.pint[3 * R.slot + 2] = ihi; }\n  }\n";
    s += "  if (DM) {\n    for (int o = 32; o > 0; o >>= 1) {\n      const int b2 = __shfl_xor(imx, o, 64);\n"
         "      imx = b2 > imx ? b2 : imx;\n    }\n    if (lane == 0) A.pmv[R.slot] = (double)imx;\n  }\n";
  } else if (red) {
    s += "  if (DS) {\n    for (int o = 32; o > 0; o >>= 1) gs += __shfl_xor(gs, o, 64);\n"
         "    if (lane == 0) A.psum[R.slot] = gs;\n  }\n";
    s += "  if (DM) {\n    for (int o = 32; o > 0; o >>= 1) {\n      const double v2 = __shfl_xor(mv, o, 64);\n"
         "      const unsigned i2 = __shfl_xor(mi, o, 64);\n      g_max_comb(mv, mi, v2, i2);\n    }\n"
         "    if (lane == 0) { A.pmv[R.slot] = mv; A.pmi[R.slot] = mi; }\n  }\n";
  }
  s += "  if (fl) atomicOr(A.counter + 1, fl);\n}\n";
  return s;
}

std::string batch_source(const GBatch &B) {
  const int vec = 16 / (int)ph::dtype_size(B.xdt);
  const int msteps = 8;
  const int sub = 64 * vec * msteps;
  std::string s = kPrelude;
  s += kGRec;
  for (int c = 0; c < (int)B.combos.size(); ++c) s += item_source(B, c, vec, msteps);
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_gbatch(GArgs A) {\n";
  s += "  const int lane = threadIdx.x & 63;\n  const int nwaves = (int)gridDim.x * 4;\n";
  s += "  int q = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);\n";
  if (B.any_slice)
    s += fmt("  __shared__ %s sx_all[4][%d];\n  %s *sx = sx_all[threadIdx.x >> 6];\n", stype(B.xdt), sub,
             stype(B.xdt));
  else
    s += fmt("  %s *sx = nullptr;\n", stype(B.xdt));
  s += "  if (A.dynamic) {\n    int t = 0;\n    if (lane == 0) t = atomicAdd(A.counter, 1);\n"
       "    q = __builtin_amdgcn_readfirstlane(t);\n  }\n";
  s += "  while (q < A.nitems) {  // every wavefront leaves once the items are exhausted\n";
  s += "    const GRec R = A.recs[q];\n    switch (R.combo) {\n";
  for (int c = 0; c < (int)B.combos.size(); ++c) {
    if (B.combos[(size_t)c].red < 0) {
      s += fmt("      case %d: item_%d<false, false>(A, R, lane, sx); break;\n", c, c);
    } else {
      s += fmt("      case %d:\n        if (R.ops == 1) item_%d<true, false>(A, R, lane, sx);\n", c, c);
      s += fmt("        else if (R.ops == 2) item_%d<false, true>(A, R, lane, sx);\n", c);
      s += fmt("        else item_%d<true, true>(A, R, lane, sx);\n        break;\n", c);
    }
  }
  s += "      default: break;\n    }\n";
  s += "    if (A.dynamic) {\n      int t = 0;\n      if (lane == 0) t = atomicAdd(A.counter, 1);\n"
       "      q = __builtin_amdgcn_readfirstlane(t);\n    } else {\n      q += nwaves;\n    }\n  }\n}\n";
  // per array (a wavefront each): lane l folds a run of consecutive items
  // in order, then the lanes' results are folded in lane order (a fixed
  // tree), so each result depends only on the item partials
  s += "extern \"C\" __global__ void __launch_bounds__(256) ph_gbatch_final(GArgs A) {\n";
  s += "  const int k = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);\n";
  s += "  if (k >= A.n) return;\n";
  s += "  const int ops = A.rops[k];\n  const int b = A.first[k], e = A.first[k + 1];\n";
  s += "  const int per = (e - b + 63) / 64;\n";
  s += "  const int lb = b + lane * per < e ? b + lane * per : e, le = lb + per < e ? lb + per : e;\n";
  s += "  if ((ops & 5) == 5 && A.sums) {  // Int32: total with the running total's lowest and highest prefix\n"
       "    long long t = 0, lo = 1LL << 61, hi = -(1LL << 61);\n"
       "    for (int i = lb; i < le; ++i) {\n"
       "      const long long l2 = t + A.pint[3 * i + 1], h2 = t + A.pint[3 * i + 2];\n"
       "      lo = l2 < lo ? l2 : lo; hi = h2 > hi ? h2 : hi; t += A.pint[3 * i];\n    }\n"
       "    for (int o = 1; o < 64; o <<= 1) {\n"
       "      const long long t2 = __shfl_down(t, o, 64), l2 = __shfl_down(lo, o, 64), h2 = __shfl_down(hi, o, 64);\n"
       "      if ((lane & (2 * o - 1)) == 0 && lane + o < 64) {\n"
       "        lo = t + l2 < lo ? t + l2 : lo; hi = t + h2 > hi ? t + h2 : hi; t += t2;\n      }\n    }\n"
       "    if (lane == 0) {\n      A.sums[k] = (double)t;\n"
       "      if (lo < -2147483648LL || hi > 2147483647LL) atomicOr(A.counter + 1, 1);\n    }\n"
       "  } else if ((ops & 1) && A.sums) {\n    double t = 0.0;\n    for (int i = lb; i < le; ++i) t += A.psum[i];\n"
       "    for (int o = 1; o < 64; o <<= 1) {\n      const double t2 = __shfl_down(t, o, 64);\n"
       "      if ((lane & (2 * o - 1)) == 0 && lane + o < 64) t += t2;\n    }\n"
       "    if (lane == 0) A.sums[k] = t;\n  }\n";
  s += "  if ((ops & 6) == 6 && A.maxes && e > b) {\n    double v = -__builtin_inf();\n"
       "    for (int i = lb; i < le; ++i) v = A.pmv[i] > v ? A.pmv[i] : v;\n"
       "    for (int o = 32; o > 0; o >>= 1) { const double v2 = __shfl_xor(v, o, 64); v = v2 > v ? v2 : v; }\n"
       "    if (lane == 0) A.maxes[k] = v;\n"
       "  } else if ((ops & 2) && A.maxes && e > b) {\n    double v = -__builtin_inf(); unsigned ix = 0xffffffffu;\n"
       "    for (int i = lb; i < le; ++i) g_max_comb(v, ix, A.pmv[i], A.pmi[i]);\n"
       "    for (int o = 32; o > 0; o >>= 1) {\n      const double v2 = __shfl_xor(v, o, 64);\n"
       "      const unsigned i2 = __shfl_xor(ix, o, 64);\n      g_max_comb(v, ix, v2, i2);\n    }\n"
       "    if (lane == 0) A.maxes[k] = v;\n  }\n}\n";
  s += "extern \"C\" __global__ void ph_gbatch_reset(GArgs A) {\n"
       "  if (threadIdx.x == 0 && blockIdx.x == 0) { A.counter[0] = 0; A.counter[1] = 0; }\n}\n";
  return s;
}

int cus() {
  static int n = 0;
  if (!n) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, ph::device()) == hipSuccess) n = p.multiProcessorCount;
    if (n <= 0) n = 256;
  }
  return n;
}

void drop_graph(GBatch *B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
}

int build_items(GBatch *B, int64_t chunk) {
  std::vector<GRec> items;
  std::vector<int32_t> first((size_t)B->n + 1);
  for (int k = 0; k < B->n; ++k) {
    first[(size_t)k] = (int32_t)items.size();
    const GArray &A = B->arrays[(size_t)k];
    if (A.combo < 0) continue;
    GRec r;
    std::memset(&r, 0, sizeof(r));
    r.off = A.off;
    r.zoff = A.zoff;
    for (int i = 0; i < 3; ++i) {
      r.shp[i] = (uint32_t)A.shp[i];
      r.f[i] = (uint32_t)A.f[i];
      r.s[i] = (uint32_t)A.s[i];
      r.cnt[i] = (uint32_t)A.cnt[i];
    }
    magic_u32(r.shp[2], &r.m_c, &r.l_c);
    magic_u32(r.shp[1], &r.m_b, &r.l_b);
    magic_u32(r.cnt[2], &r.mz_c, &r.lz_c);
    magic_u32(r.cnt[1], &r.mz_b, &r.lz_b);
    r.rev = A.rev;
    r.combo = A.combo;
    r.ops = A.rec->red >= 0 ? A.rec->red_ops : 0;
    r.scalar = A.rec->scalar;
    const bool slice = A.rec->slice && A.zn > 0;
    for (int64_t lo = 0; lo < A.n; lo += chunk) {
      r.lo = (uint32_t)lo;
      r.hi = (uint32_t)std::min(A.n, lo + chunk);
      r.zlo = slice ? (uint32_t)slice_before(A, lo) : 0u;
      r.zhi = slice ? (uint32_t)slice_before(A, r.hi) : 0u;
      r.slot = (int32_t)items.size();
      items.push_back(r);
    }
  }
  first[(size_t)B->n] = (int32_t)items.size();
  const int64_t ni = (int64_t)items.size();
  if (ni >= INT32_MAX) return ph::set_error(PH_E_ARGUMENT, "batch: too many work items (%lld)", (long long)ni);
  const std::vector<GRec> array_order = items;
  std::stable_sort(items.begin(), items.end(),
                   [](const GRec &a, const GRec &b) { return a.hi - a.lo > b.hi - b.lo; });
  PH_HIP(hipStreamSynchronize(ph::stream()));
  if (ni > B->items_cap) {
    (void)hipFree(B->d_recs);
    (void)hipFree(B->d_recs_am);
    (void)hipFree(B->d_psum);
    (void)hipFree(B->d_pmv);
    (void)hipFree(B->d_pmi);
    (void)hipFree(B->d_pint);
    B->d_pint = nullptr;
    B->d_recs = B->d_recs_am = nullptr;
    B->d_psum = B->d_pmv = nullptr;
    B->d_pmi = nullptr;
    B->items_cap = 0;
    if (hipMalloc(&B->d_recs, sizeof(GRec) * (size_t)ni) != hipSuccess ||
        hipMalloc(&B->d_recs_am, sizeof(GRec) * (size_t)ni) != hipSuccess ||
        hipMalloc(&B->d_psum, sizeof(double) * (size_t)ni) != hipSuccess ||
        hipMalloc(&B->d_pmv, sizeof(double) * (size_t)ni) != hipSuccess ||
        hipMalloc(&B->d_pmi, sizeof(unsigned) * (size_t)ni) != hipSuccess ||
        (B->any_isum && hipMalloc(&B->d_pint, 3 * sizeof(long long) * (size_t)ni) != hipSuccess))
      return ph::set_error(PH_E_OOM, "batch work items (%lld)", (long long)ni);
    B->items_cap = ni;
  }
  if (ni) {
    PH_HIP(hipMemcpy(B->d_recs, items.data(), sizeof(GRec) * (size_t)ni, hipMemcpyHostToDevice));
    PH_HIP(hipMemcpy(B->d_recs_am, array_order.data(), sizeof(GRec) * (size_t)ni, hipMemcpyHostToDevice));
  }
  B->first = first;
  PH_HIP(hipMemcpy(B->d_first, first.data(), sizeof(int32_t) * first.size(), hipMemcpyHostToDevice));
  B->nitems = ni;
  B->chunk = chunk;
  drop_graph(B);
  return PH_OK;
}

int launch_module(hipFunction_t f, unsigned grid, GArgsHost *a, const char *what) {
  size_t asz = sizeof(*a);
  void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, a, HIP_LAUNCH_PARAM_BUFFER_SIZE, &asz, HIP_LAUNCH_PARAM_END};
  PH_HIP(hipModuleLaunchKernel(f, grid, 1, 1, 256, 1, 1, 0, ph::stream(), nullptr, cfg));
  return ph::check_launch(what);
}

// The batch's kernels on the library stream (what a graph replay repeats).
int launch_batch(GBatch *B) {
  int rc;
  if (B->dynamic || B->may_flag) {
    rc = launch_module(B->k_reset, 1, &B->args, "ph_gbatch_reset");
    if (rc) return rc;
  }
  if (B->nitems) {
    rc = launch_module(B->k_main, (unsigned)B->grid, &B->args, "ph_gbatch");
    if (rc) return rc;
  }
  if (B->any_red) {
    rc = launch_module(B->k_final, (unsigned)((B->n + 3) / 4), &B->args, "ph_gbatch_final");
    if (rc) return rc;
  }
  return PH_OK;
}

// Points the launch arguments at the caller's buffers; a change of buffers
// or schedule invalidates the captured graph.
int bind(GBatch *B, const void *x, void *y, void *z, double *sums, double *maxes) {
  if (((uintptr_t)x | (uintptr_t)y) % 16 != 0)
    return ph::set_error(PH_E_ARGUMENT, "batch: x and y must be 16-byte aligned");
  bool need_y = false, need_z = false, need_s = false, need_m = false;
  for (const GArray &A : B->arrays) {
    if (A.combo < 0) continue;
    const ph_batch_rec &R = *A.rec;
    need_y = need_y || R.map >= 0;
    need_z = need_z || (R.slice && A.zn > 0);
    need_s = need_s || (R.red >= 0 && (R.red_ops & PH_RED_SUM));
    need_m = need_m || (R.red >= 0 && (R.red_ops & PH_RED_MAX));
  }
  if ((B->total > 0 && !x) || (need_y && !y) || (need_z && !z) || (need_s && !sums) || (need_m && !maxes))
    return ph::set_error(PH_E_ARGUMENT, "batch: a buffer the records write (or x) is NULL");
  GArgsHost a = B->args;
  a.recs = B->d_recs;
  a.x = x;
  a.y = y;
  a.z = z;
  a.psum = B->d_psum;
  a.pmv = B->d_pmv;
  a.pmi = B->d_pmi;
  a.first = B->d_first;
  a.rops = B->d_rops;
  a.sums = sums;
  a.maxes = maxes;
  a.counter = B->d_counter;
  a.pint = B->d_pint;
  a.nitems = (int)B->nitems;
  a.dynamic = B->dynamic;
  a.n = B->n;
  a.pad_ = 0;
  if (std::memcmp(&a, &B->args, sizeof(a)) != 0) {
    drop_graph(B);
    B->args = a;
  }
  return PH_OK;
}

int read_flags(GBatch *B, int32_t *flags_out) {
  if (!flags_out) return PH_OK;
  *flags_out = 0;
  if (!B->may_flag) return PH_OK;
  int v[2] = {0, 0};
  PH_HIP(hipMemcpyAsync(v, B->d_counter, sizeof(v), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  *flags_out = v[1];
  return PH_OK;
}

ph_view_desc contiguous(int dt, int64_t extent, int64_t off, int rank, const int64_t *shape) {
  ph_view_desc d;
  std::memset(&d, 0, sizeof(d));
  d.dtype = dt;
  d.extent = extent;
  d.view.rank = rank;
  d.view.offset = off;
  int64_t st = 1;
  for (int a = rank - 1; a >= 0; --a) {
    d.view.shape[a] = shape[a];
    d.view.stride[a] = st;
    st *= shape[a];
  }
  return d;
}

// One launch of the batch's kernel per array (its items, one wavefront
// each), then the fold of the partial results: the batch's work without the
// batching.
int run_per_array(GBatch *B) {
  int rc;
  if (B->may_flag) {
    rc = launch_module(B->k_reset, 1, &B->args, "ph_gbatch_reset");
    if (rc) return rc;
  }
  for (int k = 0; k < B->n; ++k) {
    const int32_t b = B->first[(size_t)k], e = B->first[(size_t)k + 1];
    if (e == b) continue;
    GArgsHost a = B->args;
    a.recs = B->d_recs_am + b;
    a.nitems = e - b;
    a.dynamic = 0;
    rc = launch_module(B->k_main, (unsigned)((e - b + 3) / 4), &a, "ph_gbatch (one array)");
    if (rc) return rc;
  }
  if (B->any_red) return launch_module(B->k_final, (unsigned)((B->n + 3) / 4), &B->args, "ph_gbatch_final");
  return PH_OK;
}

// The unbatched form: each record through the library's own entry points.
int run_unbatched(GBatch *B, const void *x, void *y, void *z, double *sums, double *maxes, int32_t *flags_out) {
  int32_t all = 0;
  for (int k = 0; k < B->n; ++k) {
    const GArray &A = B->arrays[(size_t)k];
    const ph_batch_rec &R = *A.rec;
    const ph_view_desc xd = contiguous(B->xdt, B->total, A.off, R.rank, R.shape);
    ph_view_desc sd = xd;
    sd.extent = B->n;
    sd.view.offset = k;
    for (int a = 0; a < R.rank; ++a) sd.view.stride[a] = 0;
    const void *ins[2] = {x, B->d_scalars};
    const ph_view_desc descs[2] = {xd, sd};
    int32_t fl = 0;
    if (R.map >= 0) {
      ph_view_desc yd = xd;
      yd.dtype = B->ydt;
      int rc = ph_map(&B->progs[(size_t)R.map], 2, ins, descs, y, &yd, &fl);
      if (rc) return rc;
      all |= fl;
    }
    if (R.red >= 0) {
      int rc = ph_reduce(&B->progs[(size_t)R.red], 2, ins, descs, -1, R.red_ops, PH_ACC_F64,
                         (R.red_ops & PH_RED_SUM) ? sums + k : nullptr, (R.red_ops & PH_RED_MAX) ? maxes + k : nullptr,
                         &fl);
      if (rc) return rc;
      all |= fl;
    }
    if (R.slice && A.zn > 0) {
      ph_view_desc rd = xd;
      int64_t cnt[3];
      for (int a = 0; a < R.rank; ++a) {
        rd.view.offset += R.first[a] * xd.view.stride[a];
        rd.view.shape[a] = R.count[a];
        rd.view.stride[a] = R.step[a] * xd.view.stride[a];
        cnt[a] = R.count[a];
      }
      const ph_view_desc zd = contiguous(B->xdt, B->ztotal, A.zoff, R.rank, cnt);
      int rc = ph_strided_copy(z, &zd, x, &rd, 0);
      if (rc) return rc;
    }
  }
  if (flags_out) *flags_out = all;
  return PH_OK;
}

}  // namespace

namespace ph {

int gbatch_compile_check(const ph_prog *prog, int x_dtype) {
  GBatch B;
  B.xdt = x_dtype;
  B.ydt = prog->out_dtype;
  B.progs.push_back(*prog);
  B.combos.push_back(Combo{0, -1});
  B.any_slice = true;
  if (prog->out_dtype == PH_F32 || prog->out_dtype == PH_F64 || prog->out_dtype == PH_I32) {
    B.combos.push_back(Combo{-1, 0});
    B.combos.push_back(Combo{0, 0});
    B.any_red = true;
  }
  return jit_compile_only(batch_source(B));
}

}  // namespace ph

extern "C" {

int ph_gbatch_destroy(void *handle);

int ph_gbatch_create(void **handle, int32_t x_dtype, int32_t y_dtype, int32_t n_progs, const ph_prog *progs,
                     int32_t n, const ph_batch_rec *recs) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (!handle || n < 0 || (n > 0 && !recs) || n_progs < 0 || n_progs > kMaxProgs || (n_progs > 0 && !progs))
    return ph::set_error(PH_E_ARGUMENT, "ph_gbatch_create arguments");
  if (x_dtype != PH_I32 && x_dtype != PH_F32 && x_dtype != PH_F64)
    return ph::set_error(PH_E_ARGUMENT, "batch: x dtype %d (I32, F32 or F64)", x_dtype);
  if (y_dtype != PH_I32 && y_dtype != PH_F32 && y_dtype != PH_F64 && y_dtype != PH_BOOL)
    return ph::set_error(PH_E_ARGUMENT, "batch: y dtype %d (I32, F32, F64 or Bool)", y_dtype);
  GBatch *B = new GBatch();
  auto fail = [&](int rc) {
    ph_gbatch_destroy(B);
    return rc;
  };
  B->xdt = x_dtype;
  B->ydt = y_dtype;
  B->n = n;
  const int in_dt[2] = {x_dtype, x_dtype};
  for (int p = 0; p < n_progs; ++p) {
    int rc = validate_prog(&progs[p], 2, in_dt);
    if (rc) return fail(rc);
    B->progs.push_back(progs[p]);
    B->may_flag = B->may_flag || prog_may_flag(&progs[p]);
  }
  B->recs.assign(recs, recs + n);
  std::map<std::pair<int, int>, int> combo_ix;
  int64_t off = 0, zoff = 0;
  for (int k = 0; k < n; ++k) {
    const ph_batch_rec &R = B->recs[(size_t)k];
    GArray a;
    std::memset(&a, 0, sizeof(a));
    a.rec = &R;
    if (R.rank < 1 || R.rank > 3) return fail(ph::set_error(PH_E_DIMENSION, "batch record %d: rank %d (1..3)", k, R.rank));
    a.rank = R.rank;
    for (int i = 0; i < 3; ++i) {
      a.shp[i] = 1;
      a.f[i] = 0;
      a.s[i] = 1;
      a.cnt[i] = 1;
    }
    a.n = 1;
    a.zn = 1;
    for (int i = 0; i < R.rank; ++i) {
      const int ax = 3 - R.rank + i;
      const int64_t d = R.shape[i];
      if (d < 1 || d > (int64_t(1) << 30))
        return fail(ph::set_error(PH_E_DIMENSION, "batch record %d: extent %lld (1..2^30)", k, (long long)d));
      a.shp[ax] = d;
      a.n *= d;
      if (R.slice) {
        const int64_t f = R.first[i], st = R.step[i], c = R.count[i];
        if (st == 0 || c < 0) return fail(ph::set_error(PH_E_ARGUMENT, "batch record %d: step 0 or negative count", k));
        if (c > 0) {
          const int64_t last = f + (c - 1) * st;
          if (f < 0 || f >= d || last < 0 || last >= d)
            return fail(ph::set_error(PH_E_INDEX, "batch record %d: region axis %d leaves [0, %lld)", k, i,
                                      (long long)d));
        }
        a.cnt[ax] = c;
        a.zn *= c;
        if (st > 0) {
          a.f[ax] = f;
          a.s[ax] = st;
        } else {  // walk the axis forwards; z's axis runs backwards
          a.f[ax] = c > 0 ? f + (c - 1) * st : 0;
          a.s[ax] = -st;
          a.rev |= 1u << (2 - ax);
        }
      }
    }
    if (a.n >= (int64_t(1) << 31))
      return fail(ph::set_error(PH_E_DIMENSION, "batch record %d: 2^31 elements or more", k));
    if (!R.slice) a.zn = 0;
    if ((R.map >= n_progs) || (R.red >= n_progs) || R.map < -1 || R.red < -1)
      return fail(ph::set_error(PH_E_ARGUMENT, "batch record %d: program index out of range", k));
    if (R.map >= 0) {
      const int odt = progs[R.map].out_dtype;
      if (odt != y_dtype)
        return fail(ph::set_error(PH_E_ARGUMENT, "batch record %d: map program gives dtype %d, y is %d", k, odt, y_dtype));
    }
    const int ops = R.red >= 0 ? R.red_ops : 0;
    if (R.red >= 0) {
      const int odt = progs[R.red].out_dtype;
      if (odt != PH_F32 && odt != PH_F64 && odt != PH_I32)
        return fail(ph::set_error(PH_E_UNSUPPORTED, "batch record %d: reductions of dtype %d (I32, F32 or F64)", k,
                                  odt));
      if (ops < 1 || ops > 3) return fail(ph::set_error(PH_E_ARGUMENT, "batch record %d: red_ops %d", k, ops));
      if (odt == PH_I32 && (ops & PH_RED_SUM)) B->any_isum = true;
    }
    const bool slice = R.slice && a.zn > 0;
    if (R.map >= 0 || R.red >= 0 || slice) {
      const auto key = std::make_pair(R.map, R.red);
      auto it = combo_ix.find(key);
      if (it == combo_ix.end()) {
        if ((int)B->combos.size() >= kMaxCombos)
          return fail(ph::set_error(PH_E_ARGUMENT, "batch: more than %d distinct (map, reduction) program pairs",
                                    kMaxCombos));
        it = combo_ix.emplace(key, (int)B->combos.size()).first;
        B->combos.push_back(Combo{R.map, R.red});
      }
      a.combo = it->second;
      B->any_red = B->any_red || R.red >= 0;
      B->any_slice = B->any_slice || slice;
    } else {
      a.combo = -1;
    }
    a.off = off;
    a.zoff = zoff;
    off += (a.n + 15) / 16 * 16;
    zoff += a.zn;
    B->arrays.push_back(a);
  }
  B->total = off;
  B->ztotal = zoff;
  B->may_flag = B->may_flag || B->any_isum;  // a checked Int32 sum can overflow
  // per-record reduction ops and scalars (the unbatched form's operand)
  std::vector<int32_t> rops((size_t)n);
  const size_t es = ph::dtype_size(x_dtype);
  std::vector<char> sc((size_t)std::max(n, 1) * es);
  for (int k = 0; k < n; ++k) {
    const ph_batch_rec &R = B->recs[(size_t)k];
    rops[(size_t)k] = R.red >= 0 ? R.red_ops | (B->progs[(size_t)R.red].out_dtype == PH_I32 ? 4 : 0) : 0;
    if (x_dtype == PH_F32) {
      const float v = (float)R.scalar;
      std::memcpy(&sc[(size_t)k * es], &v, es);
    } else if (x_dtype == PH_F64) {
      std::memcpy(&sc[(size_t)k * es], &R.scalar, es);
    } else {
      const int32_t v = (int32_t)R.scalar;
      std::memcpy(&sc[(size_t)k * es], &v, es);
    }
  }
  if (hipMalloc(&B->d_first, sizeof(int32_t) * (size_t)(n + 1)) != hipSuccess ||
      hipMalloc(&B->d_rops, sizeof(int32_t) * (size_t)std::max(n, 1)) != hipSuccess ||
      hipMalloc(&B->d_counter, 64) != hipSuccess || hipMalloc(&B->d_scalars, sc.size()) != hipSuccess)
    return fail(ph::set_error(PH_E_OOM, "batch tables"));
  if (n) {
    PH_HIP(hipMemcpy(B->d_rops, rops.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    PH_HIP(hipMemcpy(B->d_scalars, sc.data(), sc.size(), hipMemcpyHostToDevice));
  }
  PH_HIP(hipMemset(B->d_counter, 0, 64));
  B->src = batch_source(*B);
  int rc = ph::jit_kernel(B->src, "ph_gbatch", &B->k_main);
  if (!rc) rc = ph::jit_kernel2(B->src, "ph_gbatch_final", &B->k_final);
  if (!rc) rc = ph::jit_kernel2(B->src, "ph_gbatch_reset", &B->k_reset);
  if (rc) return fail(rc);
  std::memset(&B->args, 0, sizeof(B->args));
  B->grid = cus() * 2;
  B->dynamic = 0;
  rc = build_items(B, 16384);
  if (rc) return fail(rc);
  *handle = B;
  return PH_OK;
}

int ph_gbatch_sizes(void *handle, int64_t *x_elements, int64_t *z_elements, int64_t *offsets_xy,
                    int64_t *offsets_z) {
  PH_GUARD();
  GBatch *B = (GBatch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  if (x_elements) *x_elements = B->total;
  if (z_elements) *z_elements = B->ztotal;
  for (int k = 0; k < B->n; ++k) {
    if (offsets_xy) offsets_xy[k] = B->arrays[(size_t)k].off;
    if (offsets_z) offsets_z[k] = B->arrays[(size_t)k].zoff;
  }
  return PH_OK;
}

int ph_gbatch_run(void *handle, const void *x, void *y, void *z, double *sums, double *maxes, int32_t mode,
                  int32_t *flags_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  GBatch *B = (GBatch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  if (mode < 0 || mode > 3) return ph::set_error(PH_E_ARGUMENT, "batch mode %d", mode);
  int rc = bind(B, x, y, z, sums, maxes);
  if (rc) return rc;
  if (flags_out) *flags_out = 0;
  if (B->n == 0) return PH_OK;
  if (mode == 3) return run_unbatched(B, x, y, z, sums, maxes, flags_out);
  rc = ph::launch_gate("general batch");
  if (rc) return rc;
  if (mode == 2) {
    rc = run_per_array(B);
    if (rc) return rc;
    return read_flags(B, flags_out);
  }
  if (mode == 0) {
    if (!B->exec) {
      PH_HIP(hipStreamSynchronize(ph::stream()));
      PH_HIP(hipStreamBeginCapture(ph::stream(), hipStreamCaptureModeThreadLocal));
      rc = launch_batch(B);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(ph::stream(), &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return ph::set_error(PH_E_HIP, "graph capture: %s", hipGetErrorString(e));
      B->graph = g;
      PH_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    }
    PH_HIP(hipGraphLaunch(B->exec, ph::stream()));
  } else {
    rc = launch_batch(B);
    if (rc) return rc;
  }
  return read_flags(B, flags_out);
}

int ph_gbatch_tune(void *handle, const void *x, void *y, void *z, double *sums, double *maxes, int64_t *chunk_out,
                   int32_t *grid_out, int32_t *dynamic_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  GBatch *B = (GBatch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  int rc = bind(B, x, y, z, sums, maxes);
  if (rc) return rc;
  struct Sched {
    int64_t chunk;
    int grid, dynamic;
  } best{B->chunk, B->grid, B->dynamic};
  if (B->nitems) {
    hipEvent_t e0, e1;
    PH_HIP(hipEventCreate(&e0));
    PH_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    for (int64_t chunk : {1024, 2048, 4096, 8192, 32768}) {
      rc = build_items(B, chunk);
      if (rc) break;
      for (int dyn : {0, 1}) {
        for (int per_cu : {2, 4, 8}) {
          B->grid = cus() * per_cu;
          B->dynamic = dyn;
          rc = bind(B, x, y, z, sums, maxes);
          if (!rc) rc = launch_batch(B);  // warm
          if (!rc) rc = hipEventRecord(e0, ph::stream()) == hipSuccess ? PH_OK : PH_E_HIP;
          for (int r = 0; r < 3 && !rc; ++r) rc = launch_batch(B);
          if (!rc) rc = hipEventRecord(e1, ph::stream()) == hipSuccess ? PH_OK : PH_E_HIP;
          if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? PH_OK : PH_E_HIP;
          float ms = 0;
          if (!rc) rc = hipEventElapsedTime(&ms, e0, e1) == hipSuccess ? PH_OK : PH_E_HIP;
          if (rc) break;
          if (ms < best_ms) {
            best_ms = ms;
            best = Sched{chunk, B->grid, dyn};
          }
        }
        if (rc) break;
      }
      if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc == PH_E_HIP ? ph::set_error(PH_E_HIP, "batch tuning: event or launch failed") : rc;
  }
  rc = build_items(B, best.chunk);
  if (rc) return rc;
  B->grid = best.grid;
  B->dynamic = best.dynamic;
  rc = bind(B, x, y, z, sums, maxes);
  if (rc) return rc;
  if (chunk_out) *chunk_out = best.chunk;
  if (grid_out) *grid_out = best.grid;
  if (dynamic_out) *dynamic_out = best.dynamic;
  return PH_OK;
}

int ph_gbatch_destroy(void *handle) {
  PH_GUARD();
  GBatch *B = (GBatch *)handle;
  if (!B) return PH_OK;
  (void)hipStreamSynchronize(ph::stream());
  drop_graph(B);
  (void)hipFree(B->d_recs);
  (void)hipFree(B->d_recs_am);
  (void)hipFree(B->d_first);
  (void)hipFree(B->d_rops);
  (void)hipFree(B->d_counter);
  (void)hipFree(B->d_psum);
  (void)hipFree(B->d_pmv);
  (void)hipFree(B->d_pmi);
  (void)hipFree(B->d_pint);
  (void)hipFree(B->d_scalars);
  delete B;
  return PH_OK;
}

}  // extern "C"
