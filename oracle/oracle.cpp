// oracle.cpp — CPU restatement of ph-core's NArray data-parallel path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (ph-core_amd/) links,
// loads or calls this file.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker / CPU timing
// baseline.
//
// The reference (in-phase/ph-core) is Crystal and cannot be built in this
// image (no crystal compiler), so this is a single-threaded C++ restatement
// that follows the reference's iteration order and per-element operation
// sequence.  Each function cites the reference file:line it restates.
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, never -ffast-math).
//
// Pinning (how each part is checked against something other than itself):
//   range syntax / IndexRegion   — the reference's spec tables
//     (spec/spec_helper.cr:64-192, spec/index_region_spec.cr) as golden
//     fixtures, plus brute-force enumeration on tiny bounds;
//   gather / scatter / views     — spec/n_array_spec.cr:211-295,
//     spec/multi_writable_spec.cr, spec/view_util/*_spec.cr known answers,
//     plus brute force (Cartesian enumeration) on tiny shapes;
//   elementwise Int32            — spec/n_array_spec.cr:444-518 known answers,
//     int64 closed forms for the overflow boundary;
//   elementwise Float            — IEEE-754 correct rounding (numpy float32 /
//     float64 arithmetic computes the same closed form);
//   sums                         — exact sums of integer-valued data, Python
//     math.fsum on small inputs;
//   Float32 sequential sum       — numpy's sequential float32 cumsum and
//     hand-computed absorption cases;
//   batched loop                 — numpy float32 add, basic slicing
//     x[::2, ::2, ::2] and float32 cumsum, array by array;
//   heat 2-D                     — separable eigenmode decay (closed form);
//   heat 1-D example             — steady-state / symmetry invariants.
// Crystal-stdlib semantics that no reference test covers (S1-S4 below) are
// "parity unpinned": they are stated here as assumptions.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

enum { OK = 0, E_INDEX = -1, E_SHAPE = -2, E_DIMENSION = -3, E_ARGUMENT = -4,
       E_OVERFLOW = -5, E_DIVZERO = -6, E_GENERIC = -7 };

struct Err {
  int code;
  std::string msg;
};

void set_msg(char *buf, int len, const std::string &m) {
  if (buf && len > 0) {
    std::snprintf(buf, (size_t)len, "%s", m.c_str());
  }
}

int64_t sign_of(int64_t v) { return (v > 0) - (v < 0); }

// Crystal's Int#// and Int#% are floored.
int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
  return q;
}
int64_t floormod(int64_t a, int64_t b) {
  int64_t m = a % b;
  if (m != 0 && ((m < 0) != (b < 0))) m += b;
  return m;
}

}  // namespace

extern "C" {

// ===========================================================================
// Range syntax — src/range_syntax/range_syntax.cr
// ===========================================================================

// A parsed range literal: the NamedTuple parse_range returns
// (src/range_syntax/range_syntax.cr:41-69), or an Int index (kind 1).
typedef struct {
  int32_t kind;  // 0 range, 1 integer index
  int32_t has_first, has_last, has_step;
  int64_t first, last, step;
  int32_t exclusive;
  int32_t pad_;
} or_lit;

// RangeSyntax.get_size (range_syntax.cr:7-17)
static int get_size(int64_t first, int64_t last, int64_t step, int64_t *out,
                    Err *e) {
  int64_t cmp = (last > first) - (last < first);
  if (last != first && sign_of(step) != cmp) {
    e->code = E_INDEX;
    e->msg = "Could not get size of range - step direction disagrees with first and last.";
    return e->code;
  }
  if (step == 0) {  // only reachable when last == first: Crystal's `0 // 0`
    e->code = E_DIVZERO;
    e->msg = "Division by 0";
    return e->code;
  }
  if (last >= first) {
    *out = floordiv(last - first, step) + 1;
  } else {
    *out = floordiv(first - last, -step) + 1;
  }
  return OK;
}

// CoordUtil.canonicalize_index_unsafe (src/coord_util.cr:65-71)
static int64_t canon_unsafe(int64_t index, int64_t size) {
  return index < 0 ? size + index : index;
}

// RangeSyntax.infer_range (range_syntax.cr:84-136)
static int infer_range(const or_lit *lit, int64_t bound, int64_t out[4],
                       Err *e) {
  if (lit->kind == 1) {  // :84-87
    int64_t c = canon_unsafe(lit->first, bound);
    out[0] = c; out[1] = 1; out[2] = c; out[3] = 1;
    return OK;
  }
  bool has_f = lit->has_first, has_l = lit->has_last;
  int64_t step, first, temp_last;
  if (!lit->has_step) {  // :97-101
    first = has_f ? canon_unsafe(lit->first, bound) : 0;
    temp_last = has_l ? canon_unsafe(lit->last, bound) : bound - 1;
    step = (temp_last >= first) ? 1 : -1;
  } else {  // :102-105
    step = lit->step;
    first = has_f ? canon_unsafe(lit->first, bound) : (step > 0 ? 0 : bound - 1);
    temp_last = has_l ? canon_unsafe(lit->last, bound) : (step > 0 ? bound - 1 : 0);
  }
  if (has_l && lit->exclusive) {  // :108-114
    if (temp_last == first) {
      out[0] = 0; out[1] = 0; out[2] = 0; out[3] = 0;
      return OK;
    }
    temp_last -= sign_of(step);
  }
  if (first < 0 || temp_last < 0) {  // :120-122
    e->code = E_INDEX;
    e->msg = "Invalid index: At least one endpoint is negative after canonicalization.";
    return e->code;
  }
  int64_t size = 0;
  Err inner{OK, ""};
  if (get_size(first, temp_last, step, &size, &inner) != OK) {  // :124-135
    if (inner.code == E_INDEX) {
      e->code = E_INDEX;
      e->msg = "Could not canonicalize range: Conflict between implicit direction and provided step";
    } else {
      *e = inner;
    }
    return e->code;
  }
  out[0] = first;
  out[1] = step;
  out[2] = first + step * (size - 1);
  out[3] = size;
  return OK;
}

int or_infer_range(const or_lit *lit, int64_t bound, int64_t out[4], char *msg,
                   int msglen) {
  Err e{OK, ""};
  int rc = infer_range(lit, bound, out, &e);
  if (rc) set_msg(msg, msglen, e.msg);
  return rc;
}

// RangeSyntax.canonicalize_range (range_syntax.cr:138-153)
static int canonicalize_range(const or_lit *lit, int64_t bound, int64_t out[4],
                              Err *e) {
  int rc = infer_range(lit, bound, out, e);
  if (rc) return rc;
  bool valid = out[0] >= 0 && out[0] < bound && out[2] >= 0 && out[2] < bound;
  if (!valid) {
    e->code = E_INDEX;
    e->msg = "Could not canonicalize range: not a sensible index range for axis";
    return e->code;
  }
  return OK;
}

int or_canonicalize_range(const or_lit *lit, int64_t bound, int64_t out[4],
                          char *msg, int msglen) {
  Err e{OK, ""};
  int rc = canonicalize_range(lit, bound, out, &e);
  if (rc) set_msg(msg, msglen, e.msg);
  return rc;
}

// ===========================================================================
// IndexRegion — src/index_region.cr
// ===========================================================================

#define OR_MAXR 8

typedef struct {
  int32_t rank;  // proper dimensions
  int32_t drop;
  int64_t first[OR_MAXR];
  int64_t step[OR_MAXR];
  int64_t last[OR_MAXR];
  int64_t proper_shape[OR_MAXR];
  int32_t degeneracy[OR_MAXR];
  int32_t reduced_rank;
  int32_t pad_;
  int64_t reduced_shape[OR_MAXR];
} or_region;

// IndexRegion.compute_reduced_shape (index_region.cr:323-335)
static void compute_reduced(or_region *r) {
  if (r->drop) {
    int n = 0;
    for (int i = 0; i < r->rank; ++i)
      if (!r->degeneracy[i]) r->reduced_shape[n++] = r->proper_shape[i];
    if (n == 0) {
      int64_t size = 0;  // ShapeUtil.shape_to_size (src/shape_util.cr:41-50)
      if (r->rank > 0) {
        size = 1;
        for (int i = 0; i < r->rank; ++i) size *= r->proper_shape[i];
      }
      r->reduced_shape[0] = size;
      n = 1;
    }
    r->reduced_rank = n;
  } else {
    r->reduced_rank = r->rank;
    for (int i = 0; i < r->rank; ++i) r->reduced_shape[i] = r->proper_shape[i];
  }
}

// IndexRegion.new(region_literal, bound_shape, drop) (index_region.cr:192-224)
int or_region_new(const or_lit *lits, int nlit, const int64_t *bound, int rank,
                  int drop, or_region *out, char *msg, int msglen) {
  Err e{OK, ""};
  std::memset(out, 0, sizeof(*out));
  out->rank = rank;
  out->drop = drop;
  for (int i = 0; i < rank; ++i) out->step[i] = 1;
  if (nlit > rank) {
    set_msg(msg, msglen, "The region literal had more dimensions than its bound shape");
    return E_DIMENSION;
  }
  for (int i = 0; i < nlit; ++i) {
    int64_t r[4];
    if (canonicalize_range(&lits[i], bound[i], r, &e)) {
      set_msg(msg, msglen, e.msg);
      return e.code;
    }
    out->first[i] = r[0];
    out->step[i] = r[1];
    out->last[i] = r[2];
    out->proper_shape[i] = r[3];
    if (lits[i].kind == 1) out->degeneracy[i] = drop;
  }
  for (int a = nlit; a < rank; ++a) {  // :217-221
    out->last[a] = bound[a] - 1;
    out->proper_shape[a] = bound[a];
  }
  compute_reduced(out);
  return OK;
}

// RangeSyntax.ensure_nonnegative / bounded? (range_syntax.cr:19-35, 71-82)
// and IndexRegion.new(region_literal, drop) (index_region.cr:248-275)
int or_region_new_absolute(const or_lit *lits, int nlit, int drop,
                           or_region *out, char *msg, int msglen) {
  Err e{OK, ""};
  std::memset(out, 0, sizeof(*out));
  out->rank = nlit;
  out->drop = drop;
  for (int i = 0; i < nlit; ++i) {
    const or_lit *l = &lits[i];
    if (l->kind == 1) {
      if (l->first < 0) {
        set_msg(msg, msglen, "Negative indices have no meaning when a bounding shape is not provided.");
        return E_INDEX;
      }
    } else {
      if ((l->has_first && l->first < 0) || (l->has_last && l->last < 0)) {
        set_msg(msg, msglen, "Negative indices have no meaning when a bounding shape is not provided.");
        return E_INDEX;
      }
      bool bounded = true;
      if (!l->has_step || l->step >= 0) {
        if (!l->has_last) bounded = false;
      } else {
        if (!l->has_first) bounded = false;
      }
      if (!bounded) {
        set_msg(msg, msglen, "Cannot create IndexRegion without an explicit upper bound unless you provide a bounding shape");
        return E_GENERIC;
      }
    }
    if (l->kind == 1) out->degeneracy[i] = 1;
    int64_t r[4];
    if (infer_range(l, 0, r, &e)) {
      set_msg(msg, msglen, e.msg);
      return e.code;
    }
    out->first[i] = r[0];
    out->step[i] = r[1];
    out->last[i] = r[2];
    out->proper_shape[i] = r[3];
  }
  compute_reduced(out);
  return OK;
}

// IndexRegion.cover (index_region.cr:232-238)
void or_region_cover(const int64_t *shape, int rank, int drop, or_region *out) {
  std::memset(out, 0, sizeof(*out));
  out->rank = rank;
  out->drop = drop;
  for (int i = 0; i < rank; ++i) {
    out->first[i] = 0;
    out->step[i] = shape[i] == 0 ? 0 : 1;
    out->last[i] = shape[i] - 1 > 0 ? shape[i] - 1 : 0;
    out->proper_shape[i] = shape[i];
  }
  compute_reduced(out);
}

// IndexRegion.trim_axis (index_region.cr:684-703)
static void trim_axis(int64_t nb, int64_t *first, int64_t *step, int64_t *last,
                      int64_t *size) {
  if (*first >= nb) {
    if (*last >= nb) {
      *first = 0; *step = 0; *last = 0; *size = 0;
      return;
    } else if (*step < 0) {
      int64_t as = *step < 0 ? -*step : *step;
      int64_t span = (nb - 1) - *last;
      int64_t sz = floordiv(span, as) + 1;
      span -= floormod(span, as);
      *first = *last + span;
      *size = sz;
      return;
    }
  } else if (*step > 0 && *last >= nb) {
    int64_t as = *step;
    int64_t span = (nb - 1) - *first;
    span -= floormod(span, as);
    int64_t sz = floordiv(span, as) + 1;
    *last = *first + span;
    *size = sz;
    return;
  }
}

// IndexRegion#trim! (index_region.cr:502-515)
int or_region_trim(or_region *r, const int64_t *bound, int rank, char *msg,
                   int msglen) {
  if (rank != r->rank) {
    set_msg(msg, msglen, "invalid error :)");
    return E_DIMENSION;
  }
  for (int a = 0; a < rank; ++a)
    trim_axis(bound[a], &r->first[a], &r->step[a], &r->last[a], &r->proper_shape[a]);
  compute_reduced(r);
  return OK;
}

// IndexRegion.new(region_literal, bound_shape?, drop, trim_to:) (:133-168)
int or_region_new_trim(const or_lit *lits, int nlit, const int64_t *bound,
                       int drop, const int64_t *trim_to, int rank,
                       or_region *out, char *msg, int msglen) {
  Err e{OK, ""};
  std::memset(out, 0, sizeof(*out));
  out->rank = rank;
  out->drop = drop;
  for (int i = 0; i < rank; ++i) out->step[i] = 1;
  bool allow_relative = bound != nullptr;
  const int64_t *bs = bound ? bound : trim_to;
  for (int i = 0; i < nlit; ++i) {
    const or_lit *l = &lits[i];
    if (!allow_relative) {
      bool neg = l->kind == 1 ? l->first < 0
                              : ((l->has_first && l->first < 0) || (l->has_last && l->last < 0));
      if (neg) {
        set_msg(msg, msglen, "Negative indices have no meaning when a bounding shape is not provided.");
        return E_INDEX;
      }
    }
    int64_t r[4];
    if (infer_range(l, bs[i], r, &e)) {
      set_msg(msg, msglen, e.msg);
      return e.code;
    }
    out->first[i] = r[0];
    out->step[i] = r[1];
    out->last[i] = r[2];
    out->proper_shape[i] = r[3];
    if (l->kind == 1) out->degeneracy[i] = drop;
  }
  for (int a = nlit; a < rank; ++a) {
    out->last[a] = bs[a] - 1;
    out->proper_shape[a] = bs[a];
  }
  compute_reduced(out);
  return or_region_trim(out, trim_to, rank, msg, msglen);
}

// IndexRegion#reverse! (index_region.cr:533-537)
void or_region_reverse(or_region *r) {
  for (int i = 0; i < r->rank; ++i) {
    int64_t t = r->first[i];
    r->first[i] = r->last[i];
    r->last[i] = t;
    r->step[i] = -r->step[i];
  }
}

// IndexRegion#translate! (index_region.cr:577-585)
int or_region_translate(or_region *r, const int64_t *offset, int n, char *msg,
                        int msglen) {
  for (int a = 0; a < n; ++a) {
    if (offset[a] < 0 && (r->first[a] < -offset[a] || r->last[a] < -offset[a])) {
      set_msg(msg, msglen, "Can't translate to negative indices");
      return E_INDEX;
    }
  }
  for (int a = 0; a < n; ++a) {
    r->first[a] += offset[a];
    r->last[a] += offset[a];
  }
  return OK;
}

// IndexRegion#includes? (index_region.cr:443-456)
int or_region_includes(const or_region *r, const int64_t *coord, int n) {
  if (n != r->rank) return 0;
  for (int i = 0; i < n; ++i) {
    int64_t lo, hi;
    if (r->step[i] > 0) { lo = r->first[i]; hi = r->last[i]; }
    else { lo = r->last[i]; hi = r->first[i]; }
    if (coord[i] < lo || coord[i] > hi) return 0;
    if (r->step[i] == 0) return 0;  // Crystal: `x % 0` raises; an empty axis includes nothing
    if (floormod(coord[i] - r->first[i], r->step[i]) != 0) return 0;
  }
  return 1;
}

// IndexRegion#fits_in? (index_region.cr:468-478); returns 1/0, or <0 error
int or_region_fits_in(const or_region *r, const int64_t *bound, int rank,
                      char *msg, int msglen) {
  if (rank != r->rank) {
    set_msg(msg, msglen, "The bound shape provided had a different number of dimensions than this IndexRegion, so fits_in? is meaningless.");
    return E_DIMENSION;
  }
  for (int i = 0; i < rank; ++i) {
    int64_t m = r->first[i] > r->last[i] ? r->first[i] : r->last[i];
    if (bound[i] <= m) return 0;
  }
  return 1;
}

// IndexRegion#local_to_absolute_unsafe (index_region.cr:621-637)
void or_region_local_to_absolute(const or_region *r, const int64_t *local,
                                 int64_t *out) {
  if (r->drop) {
    int la = 0;
    for (int i = 0; i < r->rank; ++i) {
      if (r->degeneracy[i]) out[i] = r->first[i];
      else { out[i] = r->first[i] + local[la] * r->step[i]; ++la; }
    }
  } else {
    for (int i = 0; i < r->rank; ++i) out[i] = r->first[i] + local[i] * r->step[i];
  }
}

// Buffered.axis_strides (src/buffered/buffered.cr:15-24)
static void axis_strides(const int64_t *shape, int rank, int64_t *st) {
  if (rank == 0) return;
  st[rank - 1] = 1;
  for (int i = rank - 2; i >= 0; --i) st[i] = st[i + 1] * shape[i + 1];
}

static int64_t region_size(const or_region *r) {
  int64_t s = 1;
  for (int i = 0; i < r->rank; ++i) s *= r->proper_shape[i];
  return r->rank ? s : 0;
}

// NArray#unsafe_fetch_chunk (src/n_array.cr:450-453): the buffered lex
// odometer walks the region's proper coordinates and keeps the flat buffer
// index up to date incrementally: its start index is
// src/buffered/indexed/stride_iterator.cr:10-13, its step
// src/buffered/indexed/lex_iterator.cr:7-21.
int64_t or_gather(const void *src, const int64_t *shape, int rank, int esize,
                  const or_region *r, void *out) {
  int64_t n = region_size(r);
  if (n == 0) return 0;
  for (int i = 0; i < r->rank; ++i)
    if (r->step[i] == 0) return 0;  // StrideIterator#next: empty if any step is 0
  int64_t bstep[OR_MAXR];
  axis_strides(shape, rank, bstep);
  int64_t coord[OR_MAXR];
  int64_t idx = 0;
  for (int i = 0; i < rank; ++i) {
    coord[i] = r->first[i];
    idx += bstep[i] * r->first[i];
  }
  const char *s = (const char *)src;
  char *o = (char *)out;
  int64_t k = 0;
  for (;;) {
    std::memcpy(o + k * esize, s + idx * esize, (size_t)esize);
    ++k;
    int i = rank - 1;
    for (; i >= 0; --i) {
      if (coord[i] == r->last[i]) {
        idx -= (coord[i] - r->first[i]) * bstep[i];
        coord[i] = r->first[i];
        if (i == 0) return k;
      } else {
        coord[i] += r->step[i];
        idx += bstep[i] * r->step[i];
        break;
      }
    }
  }
}

// NArray#unsafe_set_chunk(region, src) (src/n_array.cr:484-492): src
// elements in lex order, destination by the same odometer.
int64_t or_scatter(void *dst, const int64_t *shape, int rank, int esize,
                   const or_region *r, const void *src, int64_t src_count) {
  int64_t n = region_size(r);
  if (n == 0 || src_count == 0) return 0;
  for (int i = 0; i < r->rank; ++i)
    if (r->step[i] == 0) return 0;
  int64_t bstep[OR_MAXR];
  axis_strides(shape, rank, bstep);
  int64_t coord[OR_MAXR];
  int64_t idx = 0;
  for (int i = 0; i < rank; ++i) {
    coord[i] = r->first[i];
    idx += bstep[i] * r->first[i];
  }
  char *d = (char *)dst;
  const char *s = (const char *)src;
  int64_t k = 0;
  while (k < src_count) {
    std::memcpy(d + idx * esize, s + k * esize, (size_t)esize);
    ++k;
    int i = rank - 1;
    for (; i >= 0; --i) {
      if (coord[i] == r->last[i]) {
        idx -= (coord[i] - r->first[i]) * bstep[i];
        coord[i] = r->first[i];
        if (i == 0) return k;
      } else {
        coord[i] += r->step[i];
        idx += bstep[i] * r->step[i];
        break;
      }
    }
  }
  return k;
}

// NArray#unsafe_set_chunk(region, value) (src/n_array.cr:495-500)
int64_t or_fill_region(void *dst, const int64_t *shape, int rank, int esize,
                       const or_region *r, const void *value) {
  int64_t n = region_size(r);
  if (n == 0) return 0;
  std::vector<char> rep((size_t)(n * esize));
  for (int64_t k = 0; k < n; ++k) std::memcpy(&rep[(size_t)(k * esize)], value, (size_t)esize);
  return or_scatter(dst, shape, rank, esize, r, rep.data(), n);
}

// ===========================================================================
// View transforms — src/view_util/transforms.cr, src/view.cr
// ===========================================================================

typedef struct {
  int32_t kind;  // 0 region, 1 reshape, 2 permute, 3 reverse
  int32_t rank_a;
  int64_t a[OR_MAXR];  // reshape: src_shape; permute: pattern; reverse: shape
  int32_t rank_b;
  int32_t pad_;
  int64_t b[OR_MAXR];  // reshape: new_shape
  or_region region;    // kind 0
} or_transform;

// One transform's apply (src/view_util/transforms.cr:215-221 region,
// :184-188 reshape, :267-270 permute, :298-304 reverse).
static int apply_transform(const or_transform *t, const int64_t *in, int nin,
                           int64_t *out, int *nout) {
  switch (t->kind) {
    case 0: {
      or_region_local_to_absolute(&t->region, in, out);
      *nout = t->region.rank;
      return OK;
    }
    case 1: {  // ReshapeTransform: linearise in new_shape, delinearise in src_shape
      int64_t vs[OR_MAXR];
      axis_strides(t->b, t->rank_b, vs);
      int64_t index = 0;
      for (int i = 0; i < nin; ++i) index += in[i] * vs[i];
      for (int d = t->rank_a - 1; d >= 0; --d) {
        out[d] = floormod(index, t->a[d]);
        index = floordiv(index, t->a[d]);
      }
      *nout = t->rank_a;
      return OK;
    }
    case 2: {  // PermuteTransform#unpermute: src[pattern[i]] = view[i]
      for (int i = 0; i < t->rank_a; ++i) out[i] = in[i];
      for (int i = 0; i < t->rank_a; ++i) out[t->a[i]] = in[i];
      *nout = t->rank_a;
      return OK;
    }
    case 3: {  // ReverseTransform
      for (int i = 0; i < t->rank_a; ++i) out[i] = t->a[i] - 1 - in[i];
      *nout = t->rank_a;
      return OK;
    }
  }
  return E_ARGUMENT;
}

// ComposedTransform#apply (transforms.cr:77-91): transforms are stored newest
// first (compose! unshifts, :46-48) and applied in that order, taking a view
// coordinate to a source coordinate.
static void chain_apply(const or_transform *chain, int nchain, const int64_t *in,
                        int nin, int64_t *out, int *nout) {
  int64_t a[OR_MAXR], b[OR_MAXR];
  int na = nin;
  for (int i = 0; i < nin; ++i) a[i] = in[i];
  for (int t = 0; t < nchain; ++t) {
    int nb = 0;
    apply_transform(&chain[t], a, na, b, &nb);
    for (int i = 0; i < nb; ++i) a[i] = b[i];
    na = nb;
  }
  for (int i = 0; i < na; ++i) out[i] = a[i];
  *nout = na;
}

int or_chain_apply(const or_transform *chain, int nchain, const int64_t *in,
                   int nin, int64_t *out) {
  int n = 0;
  chain_apply(chain, nchain, in, nin, out, &n);
  return n;
}

// View#to_narr (src/view.cr:123-126) = NArray.build over the view's shape in
// lex order, each element `src.unsafe_fetch_element(transform.apply(coord))`
// (src/view.cr:109-117, src/n_array.cr:456-458).
int64_t or_view_gather(const void *src, const int64_t *src_shape, int src_rank,
                       int esize, const or_transform *chain, int nchain,
                       const int64_t *view_shape, int view_rank, void *out) {
  int64_t n = 1;
  for (int i = 0; i < view_rank; ++i) n *= view_shape[i];
  if (view_rank == 0 || n == 0) return 0;
  int64_t bs[OR_MAXR];
  axis_strides(src_shape, src_rank, bs);
  int64_t coord[OR_MAXR] = {0};
  const char *s = (const char *)src;
  char *o = (char *)out;
  for (int64_t k = 0; k < n; ++k) {
    int64_t sc[OR_MAXR];
    int ns = 0;
    chain_apply(chain, nchain, coord, view_rank, sc, &ns);
    int64_t idx = 0;
    for (int i = 0; i < ns; ++i) idx += sc[i] * bs[i];
    std::memcpy(o + k * esize, s + idx * esize, (size_t)esize);
    for (int i = view_rank - 1; i >= 0; --i) {
      if (++coord[i] < view_shape[i]) break;
      coord[i] = 0;
    }
  }
  return n;
}

// MultiWritable#unsafe_set_chunk over a MutableView (src/multi_writable.cr:
// 29-36, src/mutable_view.cr:16-18): region coordinates in lex order paired
// with src elements in lex order, each written through the transform chain.
// `r` is the region in the view's coordinates (MultiWritable#set_chunk,
// src/multi_writable.cr:55-63).
int64_t or_view_scatter(void *dst, const int64_t *dst_shape, int dst_rank,
                        int esize, const or_transform *chain, int nchain,
                        const or_region *r, const void *src, int64_t src_count) {
  int64_t n = region_size(r);
  if (n == 0) return 0;
  for (int i = 0; i < r->rank; ++i)
    if (r->step[i] == 0) return 0;
  int64_t bs[OR_MAXR];
  axis_strides(dst_shape, dst_rank, bs);
  int64_t coord[OR_MAXR];
  for (int i = 0; i < r->rank; ++i) coord[i] = r->first[i];
  char *d = (char *)dst;
  const char *s = (const char *)src;
  int64_t k = 0;
  while (k < src_count) {
    int64_t sc[OR_MAXR];
    int ns = 0;
    chain_apply(chain, nchain, coord, r->rank, sc, &ns);
    int64_t idx = 0;
    for (int i = 0; i < ns; ++i) idx += sc[i] * bs[i];
    std::memcpy(d + idx * esize, s + k * esize, (size_t)esize);
    ++k;
    int i = r->rank - 1;
    for (; i >= 0; --i) {
      if (coord[i] == r->last[i]) {
        coord[i] = r->first[i];
        if (i == 0) return k;
      } else {
        coord[i] += r->step[i];
        break;
      }
    }
  }
  return k;
}

// NArray#[]=(mask, value) (src/n_array.cr:510-519): walk the mask's buffer
// in index order and store `value` where it is true; and the form with a
// source array (:523-536), whose lexicographic walk over a row-major mask,
// source and destination of one shape is the same buffer index order.
void or_set_mask(void *dst, int64_t n, int esize, const unsigned char *mask, const void *src,
                 int src_is_array) {
  unsigned char *d = (unsigned char *)dst;
  const unsigned char *s = (const unsigned char *)src;
  for (int64_t idx = 0; idx < n; ++idx)
    if (mask[idx]) std::memcpy(d + idx * esize, s + (src_is_array ? idx * esize : 0), (size_t)esize);
}

// MultiIndexable#tile (src/multi_indexable.cr:818-827) with TilingLexIterator
// (src/multi_indexable/tiling_lex_iterator.cr:27-45): out[c] = src[c mod shape].
int64_t or_tile(const void *src, const int64_t *shape, int rank, int esize,
                const int64_t *counts, void *out) {
  int64_t ns[OR_MAXR];
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) {
    ns[i] = shape[i] * counts[i];
    n *= ns[i];
  }
  if (rank == 0 || n == 0) return 0;
  int64_t bs[OR_MAXR];
  axis_strides(shape, rank, bs);
  int64_t coord[OR_MAXR] = {0};
  const char *s = (const char *)src;
  char *o = (char *)out;
  for (int64_t k = 0; k < n; ++k) {
    int64_t idx = 0;
    for (int i = 0; i < rank; ++i) idx += floormod(coord[i], shape[i]) * bs[i];
    std::memcpy(o + k * esize, s + idx * esize, (size_t)esize);
    for (int i = rank - 1; i >= 0; --i) {
      if (++coord[i] < ns[i]) break;
      coord[i] = 0;
    }
  }
  return n;
}

// ===========================================================================
// Elementwise arithmetic under Crystal semantics.
//   The operator macros (src/multi_indexable.cr:931-985) call Crystal's own
//   number operators, which live in Crystal's stdlib (1.0.0, shard.yml:8),
//   not in the reference.  Assumptions (parity unpinned except Int32 + - * **
//   which spec/n_array_spec.cr:444-518 covers):
//   S1 Float32 op Float64 -> Float64; Float op Int -> that Float; Int op
//      Float -> that Float.
//   S2 Int32 + - * ** and unary - raise OverflowError on overflow; &+ &- &*
//      &** wrap; Int32 / Int32 -> Float64; // and % are floored and raise
//      DivisionByZeroError on 0; ** with a negative exponent raises
//      ArgumentError.
//   S3 Float ** Int32 is llvm.powi: binary powering by repeated squaring.
//   S4 Float // is floor(a/b); Float % is a - b*floor(a/b).
// ===========================================================================

enum { DT_I32 = 0, DT_F32 = 1, DT_F64 = 2, DT_BF16 = 3, DT_BOOL = 4 };

struct Val {
  int dt;
  int64_t i;  // I32 / BOOL
  double f;   // F32 (exactly representable) / F64
};

static double bf16_to_f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return (double)f;
}

static Val load_val(int dt, const void *p, int64_t k) {
  Val v{dt, 0, 0.0};
  switch (dt) {
    case DT_I32: v.i = ((const int32_t *)p)[k]; break;
    case DT_F32: v.f = (double)((const float *)p)[k]; break;
    case DT_F64: v.f = ((const double *)p)[k]; break;
    case DT_BF16: v.dt = DT_F32; v.f = bf16_to_f(((const uint16_t *)p)[k]); break;
    case DT_BOOL: v.i = ((const uint8_t *)p)[k] ? 1 : 0; break;
  }
  return v;
}

static void store_val(const Val &v, void *p, int64_t k) {
  switch (v.dt) {
    case DT_I32: ((int32_t *)p)[k] = (int32_t)v.i; break;
    case DT_F32: ((float *)p)[k] = (float)v.f; break;
    case DT_F64: ((double *)p)[k] = v.f; break;
    case DT_BOOL: ((uint8_t *)p)[k] = (uint8_t)(v.i ? 1 : 0); break;
  }
}

static bool is_float(int dt) { return dt == DT_F32 || dt == DT_F64; }

static int arith_type(int a, int b) {
  if (a == DT_I32 && b == DT_I32) return DT_I32;
  if (a == DT_F64 || b == DT_F64) return DT_F64;
  return DT_F32;
}

static double as_float(const Val &v, int dt) {
  double x = is_float(v.dt) ? v.f : (double)v.i;
  if (dt == DT_F32) return (double)(float)x;
  return x;
}

static int32_t wrap32(int64_t x) { return (int32_t)(uint32_t)(uint64_t)x; }

static bool fits32(int64_t x) { return x >= INT32_MIN && x <= INT32_MAX; }

// Float ** Int32 by binary powering (compiler-rt __powidf2 / __powisf2
// algorithm: r *= a on set bits, a *= a between bits, 1/r for negative b).
static double powi_d(double a, int32_t b) {
  bool recip = b < 0;
  int64_t e = b;
  if (e < 0) e = -e;
  double r = 1.0;
  for (;;) {
    if (e & 1) r *= a;
    e /= 2;
    if (e == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}
static float powi_f(float a, int32_t b) {
  bool recip = b < 0;
  int64_t e = b;
  if (e < 0) e = -e;
  float r = 1.0f;
  for (;;) {
    if (e & 1) r *= a;
    e /= 2;
    if (e == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

// Crystal Int#** (checked or wrapping): square-and-multiply with the square
// skipped after the last bit.
static int32_t ipow(int32_t base, int32_t e, bool checked, int *flags) {
  if (e < 0) {
    *flags |= 4;
    return 0;
  }
  int64_t result = 1, k = base;
  int64_t ex = e;
  while (ex > 0) {
    if (ex & 1) {
      int64_t t = result * k;
      if (checked && !fits32(t)) *flags |= 1;
      result = wrap32(t);
    }
    ex >>= 1;
    if (ex > 0) {
      int64_t t = k * k;
      if (checked && !fits32(t)) *flags |= 1;
      k = wrap32(t);
    }
  }
  return (int32_t)result;
}

enum {
  OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_FDIV, OP_MOD, OP_POW, OP_WADD, OP_WSUB,
  OP_WMUL, OP_WPOW, OP_AND, OP_OR, OP_XOR, OP_GT, OP_LT, OP_GE, OP_LE, OP_EQ,
  OP_NE, OP_CMP, OP_NEG, OP_POS, OP_BNOT, OP_SQRT, OP_ABS
};

static Val fval(int dt, double x) {
  Val v{dt, 0, dt == DT_F32 ? (double)(float)x : x};
  return v;
}

static Val binop(int op, const Val &a, const Val &b, int *flags) {
  Val r{DT_I32, 0, 0.0};
  switch (op) {
    case OP_ADD: case OP_SUB: case OP_MUL: {
      int t = arith_type(a.dt, b.dt);
      if (t == DT_I32) {
        int64_t x = op == OP_ADD ? a.i + b.i : op == OP_SUB ? a.i - b.i : a.i * b.i;
        if (!fits32(x)) *flags |= 1;
        r.dt = DT_I32; r.i = wrap32(x);
        return r;
      }
      if (t == DT_F32) {
        float x = (float)as_float(a, t), y = (float)as_float(b, t);
        float z = op == OP_ADD ? x + y : op == OP_SUB ? x - y : x * y;
        return fval(DT_F32, z);
      }
      double x = as_float(a, t), y = as_float(b, t);
      double z = op == OP_ADD ? x + y : op == OP_SUB ? x - y : x * y;
      return fval(DT_F64, z);
    }
    case OP_WADD: case OP_WSUB: case OP_WMUL: {
      int64_t x = op == OP_WADD ? a.i + b.i : op == OP_WSUB ? a.i - b.i : a.i * b.i;
      r.dt = DT_I32; r.i = wrap32(x);
      return r;
    }
    case OP_DIV: {
      int t = arith_type(a.dt, b.dt);
      if (t == DT_I32) t = DT_F64;  // Int32 / Int32 -> Float64
      if (t == DT_F32) {
        float z = (float)as_float(a, t) / (float)as_float(b, t);
        return fval(DT_F32, z);
      }
      return fval(DT_F64, as_float(a, t) / as_float(b, t));
    }
    case OP_FDIV: case OP_MOD: {
      int t = arith_type(a.dt, b.dt);
      if (t == DT_I32) {
        if (b.i == 0) { *flags |= 2; r.dt = DT_I32; r.i = 0; return r; }
        if (op == OP_FDIV) {
          int64_t q = floordiv(a.i, b.i);
          if (!fits32(q)) *flags |= 1;
          r.dt = DT_I32; r.i = wrap32(q);
        } else {
          r.dt = DT_I32; r.i = floormod(a.i, b.i);
        }
        return r;
      }
      if (t == DT_F32) {
        float x = (float)as_float(a, t), y = (float)as_float(b, t);
        float q = std::floor(x / y);
        if (op == OP_FDIV) return fval(DT_F32, q);
        float m = y * q;
        return fval(DT_F32, x - m);
      }
      double x = as_float(a, t), y = as_float(b, t);
      double q = std::floor(x / y);
      if (op == OP_FDIV) return fval(DT_F64, q);
      double m = y * q;
      return fval(DT_F64, x - m);
    }
    case OP_POW: case OP_WPOW: {
      if (a.dt == DT_I32 && b.dt == DT_I32) {
        r.dt = DT_I32;
        r.i = ipow((int32_t)a.i, (int32_t)b.i, op == OP_POW, flags);
        return r;
      }
      if (is_float(a.dt) && b.dt == DT_I32) {
        if (a.dt == DT_F32) return fval(DT_F32, powi_f((float)a.f, (int32_t)b.i));
        return fval(DT_F64, powi_d(a.f, (int32_t)b.i));
      }
      int t = arith_type(a.dt, b.dt);
      if (t == DT_F32 && a.dt == DT_F32)
        return fval(DT_F32, std::pow((float)as_float(a, t), (float)as_float(b, t)));
      return fval(DT_F64, std::pow(as_float(a, DT_F64), as_float(b, DT_F64)));
    }
    case OP_AND: case OP_OR: case OP_XOR: {
      int64_t x = op == OP_AND ? (a.i & b.i) : op == OP_OR ? (a.i | b.i) : (a.i ^ b.i);
      r.dt = (a.dt == DT_BOOL && b.dt == DT_BOOL) ? DT_BOOL : DT_I32;
      r.i = r.dt == DT_BOOL ? (x & 1) : wrap32(x);
      return r;
    }
    case OP_GT: case OP_LT: case OP_GE: case OP_LE: case OP_EQ: case OP_NE:
    case OP_CMP: {
      int c;
      bool unordered = false;
      if ((a.dt == DT_I32 && b.dt == DT_I32) || (a.dt == DT_BOOL && b.dt == DT_BOOL)) {
        c = (a.i > b.i) - (a.i < b.i);
      } else {
        double x = is_float(a.dt) ? a.f : (double)a.i;
        double y = is_float(b.dt) ? b.f : (double)b.i;
        unordered = std::isnan(x) || std::isnan(y);
        c = (x > y) - (x < y);
      }
      if (op == OP_CMP) { r.dt = DT_I32; r.i = c; return r; }
      r.dt = DT_BOOL;
      bool v;
      switch (op) {
        case OP_GT: v = !unordered && c > 0; break;
        case OP_LT: v = !unordered && c < 0; break;
        case OP_GE: v = !unordered && c >= 0; break;
        case OP_LE: v = !unordered && c <= 0; break;
        case OP_EQ: v = !unordered && c == 0; break;
        default: v = unordered || c != 0; break;
      }
      r.i = v ? 1 : 0;
      return r;
    }
  }
  *flags |= 0x100;
  return r;
}

static Val unop(int op, const Val &a, int *flags) {
  Val r = a;
  switch (op) {
    case OP_NEG:
      if (a.dt == DT_I32) {
        if (a.i == INT32_MIN) *flags |= 1;
        r.i = wrap32(-a.i);
      } else {
        r.f = -a.f;
      }
      return r;
    case OP_POS: return r;
    case OP_BNOT: r.i = wrap32(~a.i); return r;
    case OP_SQRT:
      if (a.dt == DT_F32) return fval(DT_F32, std::sqrt((float)a.f));
      return fval(DT_F64, std::sqrt(is_float(a.dt) ? a.f : (double)a.i));
    case OP_ABS:
      if (a.dt == DT_I32) {
        if (a.i == INT32_MIN) *flags |= 1;
        r.i = wrap32(a.i < 0 ? -a.i : a.i);
      } else {
        r.f = std::fabs(a.f);
      }
      return r;
  }
  *flags |= 0x100;
  return r;
}

// Elementwise binary op over n elements (a_scalar / b_scalar: that operand is
// one value broadcast, the `narr op scalar` / `scalar op narr` forms,
// src/multi_indexable.cr:949-951, src/patches/number.cr:5-15).  Returns the
// result dtype; *flags gets the OR of raised conditions.
int or_binop(int op, int dta, const void *a, int a_scalar, int dtb,
             const void *b, int b_scalar, int64_t n, void *out, int *flags) {
  int rdt = -1;
  for (int64_t k = 0; k < n; ++k) {
    Val x = load_val(dta, a, a_scalar ? 0 : k);
    Val y = load_val(dtb, b, b_scalar ? 0 : k);
    Val z = binop(op, x, y, flags);
    store_val(z, out, k);
    rdt = z.dt;
  }
  if (n == 0) {
    Val x = {dta == DT_BF16 ? DT_F32 : dta, 1, 1.0}, y = {dtb == DT_BF16 ? DT_F32 : dtb, 1, 1.0};
    int f = 0;
    rdt = binop(op, x, y, &f).dt;
  }
  return rdt;
}

int or_unop(int op, int dta, const void *a, int64_t n, void *out, int *flags) {
  int rdt = dta == DT_BF16 ? DT_F32 : dta;
  for (int64_t k = 0; k < n; ++k) {
    Val z = unop(op, load_val(dta, a, k), flags);
    store_val(z, out, k);
    rdt = z.dt;
  }
  return rdt;
}

// --- Expression evaluation: a closure given as a prefix tree ---------------
//   expr := "x<k>" | <int>"i" | <float>"f" | <float>"d" | "true" | "false"
//         | "(" op expr [expr] ")"
//   op   := add sub mul div fdiv mod pow wadd wsub wmul wpow and or xor gt lt
//           ge le eq ne cmp neg pos bnot sqrt abs
struct Node {
  int kind;  // 0 input, 1 const, 2 unary, 3 binary
  int op;
  int input;
  Val c;
  int l, r;
};

struct Parser {
  const char *s = nullptr;
  std::vector<Node> nodes;
  bool ok = true;
  void ws() { while (*s == ' ' || *s == '\n' || *s == '\t') ++s; }
  int parse() {
    ws();
    if (*s == '(') {
      ++s;
      ws();
      const char *b = s;
      while (*s && *s != ' ' && *s != ')') ++s;
      std::string name(b, s);
      static const char *names[] = {"add", "sub", "mul", "div", "fdiv", "mod", "pow",
                                    "wadd", "wsub", "wmul", "wpow", "and", "or", "xor",
                                    "gt", "lt", "ge", "le", "eq", "ne", "cmp", "neg",
                                    "pos", "bnot", "sqrt", "abs"};
      int op = -1;
      for (int i = 0; i < (int)(sizeof(names) / sizeof(names[0])); ++i)
        if (name == names[i]) op = i;
      if (op < 0) { ok = false; return -1; }
      int l = parse();
      ws();
      Node n{};
      n.op = op;
      n.l = l;
      if (*s == ')') {
        n.kind = 2;
      } else {
        n.kind = 3;
        n.r = parse();
        ws();
      }
      if (*s != ')') { ok = false; return -1; }
      ++s;
      nodes.push_back(n);
      return (int)nodes.size() - 1;
    }
    const char *b = s;
    while (*s && *s != ' ' && *s != ')' && *s != '(') ++s;
    std::string tok(b, s);
    Node n{};
    if (tok.size() >= 2 && tok[0] == 'x') {
      n.kind = 0;
      n.input = std::atoi(tok.c_str() + 1);
    } else if (tok == "true" || tok == "false") {
      n.kind = 1;
      n.c = Val{DT_BOOL, tok == "true" ? 1 : 0, 0.0};
    } else if (!tok.empty()) {
      char suf = tok.back();
      std::string num = tok.substr(0, tok.size() - 1);
      n.kind = 1;
      if (suf == 'i') n.c = Val{DT_I32, std::strtoll(num.c_str(), nullptr, 10), 0.0};
      else if (suf == 'f') n.c = fval(DT_F32, (double)std::strtof(num.c_str(), nullptr));
      else if (suf == 'd') n.c = Val{DT_F64, 0, std::strtod(num.c_str(), nullptr)};
      else { ok = false; return -1; }
    } else {
      ok = false;
      return -1;
    }
    nodes.push_back(n);
    return (int)nodes.size() - 1;
  }
};

static Val eval_node(const std::vector<Node> &nodes, int id, const Val *in, int *flags) {
  const Node &n = nodes[(size_t)id];
  switch (n.kind) {
    case 0: return in[n.input];
    case 1: return n.c;
    case 2: return unop(n.op, eval_node(nodes, n.l, in, flags), flags);
    default: {
      Val a = eval_node(nodes, n.l, in, flags);
      Val b = eval_node(nodes, n.r, in, flags);
      return binop(n.op, a, b, flags);
    }
  }
}

// Evaluates `expr` elementwise over n_in inputs of n elements each (the
// map_with zip of src/multi_indexable.cr:1080-1102, or a ProcView's proc over
// its source, src/proc_view.cr:28-30).  Returns the result dtype, or -1 on
// a parse error.
int or_eval(const char *expr, int n_in, const int *in_dtypes,
            const void *const *ins, int64_t n, void *out, int *flags) {
  Parser p;
  p.s = expr;
  int root = p.parse();
  if (!p.ok) return -1;
  int rdt = -1;
  Val vals[8];
  for (int64_t k = 0; k < n; ++k) {
    for (int j = 0; j < n_in; ++j) vals[j] = load_val(in_dtypes[j], ins[j], k);
    Val z = eval_node(p.nodes, root, vals, flags);
    store_val(z, out, k);
    rdt = z.dt;
  }
  if (n == 0) {
    for (int j = 0; j < n_in; ++j) vals[j] = Val{in_dtypes[j] == DT_BF16 ? DT_F32 : in_dtypes[j], 1, 1.0};
    int f = 0;
    rdt = eval_node(p.nodes, root, vals, &f).dt;
  }
  return rdt;
}

// ===========================================================================
// Lazy sources: the reference reads a source or mask element only when its
// loop reaches it, so a source that is a view (or a ProcView) of the array
// being written sees the writes of the elements before it.
// ===========================================================================

// One operand of a lazily evaluated expression: a buffer (read at every use,
// so writes made since are seen), its row-major shape, and the transform
// chain (newest first) taking an expression coordinate to a buffer
// coordinate.  An empty chain means the coordinate is the buffer's own.
typedef struct {
  const void *buf;
  int32_t dtype;
  int32_t rank;
  int64_t shape[OR_MAXR];
  const or_transform *chain;
  int32_t nchain;
  int32_t pad_;
} or_input;

static Val read_input(const or_input &in, const int64_t *coord, int rank) {
  int64_t sc[OR_MAXR];
  int ns = 0;
  chain_apply(in.chain, in.nchain, coord, rank, sc, &ns);
  int64_t bs[OR_MAXR];
  axis_strides(in.shape, in.rank, bs);
  int64_t idx = 0;
  for (int i = 0; i < ns; ++i) idx += sc[i] * bs[i];
  return load_val(in.dtype, in.buf, idx);
}

static Val eval_lazy(const Parser &p, int root, int n_in, const or_input *ins, const int64_t *coord,
                     int rank, int *flags) {
  Val vals[8];
  for (int j = 0; j < n_in && j < 8; ++j) vals[j] = read_input(ins[j], coord, rank);
  return eval_node(p.nodes, root, vals, flags);
}

// NArray#unsafe_set_chunk(region, src) (src/n_array.cr:484-492) and, with a
// non-empty `dchain`, MultiWritable#unsafe_set_chunk through a MutableView
// (src/multi_writable.cr:29-36, src/mutable_view.cr:16-18), where `src` is
// the lazily evaluated `expr` over `ins` with shape src_shape (a View is the
// expression "x0", a ProcView its block): element k of src (lexicographic
// over src_shape, `src.each`) is evaluated when the loop reaches it and
// written to the k-th coordinate of the region.  Returns the count written,
// or -1 if `expr` does not parse.
int64_t or_set_chunk_lazy(void *dst, const int64_t *dst_shape, int dst_rank, const or_transform *dchain,
                          int ndchain, const or_region *r, const char *expr, int n_in, const or_input *ins,
                          const int64_t *src_shape, int src_rank, int *flags) {
  Parser p;
  p.s = expr;
  const int root = p.parse();
  if (!p.ok) return -1;
  int64_t n = region_size(r);
  if (n == 0) return 0;
  for (int i = 0; i < r->rank; ++i)
    if (r->step[i] == 0) return 0;
  int64_t bs[OR_MAXR];
  axis_strides(dst_shape, dst_rank, bs);
  int64_t coord[OR_MAXR], sc[OR_MAXR] = {0};
  for (int i = 0; i < r->rank; ++i) coord[i] = r->first[i];
  int64_t k = 0;
  for (;;) {
    const Val v = eval_lazy(p, root, n_in, ins, sc, src_rank, flags);
    int64_t dc[OR_MAXR];
    int nd = 0;
    chain_apply(dchain, ndchain, coord, r->rank, dc, &nd);
    int64_t idx = 0;
    for (int i = 0; i < nd; ++i) idx += dc[i] * bs[i];
    store_val(v, dst, idx);
    ++k;
    for (int i = src_rank - 1; i >= 0; --i) {
      if (++sc[i] < src_shape[i]) break;
      sc[i] = 0;
    }
    int i = r->rank - 1;
    for (; i >= 0; --i) {
      if (coord[i] == r->last[i]) {
        coord[i] = r->first[i];
        if (i == 0) return k;
      } else {
        coord[i] += r->step[i];
        break;
      }
    }
  }
}

// NArray#[]=(mask, value) (src/n_array.cr:523-551): for every coordinate of
// the array in lexicographic order, `mask.get(coord)` is evaluated, and where
// it holds, `value.get(coord)` (or the scalar) is written at that
// coordinate.  `mask_expr` and `value_expr` are lazily evaluated expressions
// over `ins` (value_expr NULL: the scalar `*scalar` of the array's dtype).
// Returns 0, or -1 if an expression does not parse.
int or_set_mask_lazy(void *dst, int dtype, const int64_t *shape, int rank, const char *mask_expr,
                     const char *value_expr, const void *scalar, int n_in, const or_input *ins, int *flags) {
  Parser pm;
  pm.s = mask_expr;
  const int mroot = pm.parse();
  if (!pm.ok) return -1;
  Parser pv;
  int vroot = -1;
  if (value_expr) {
    pv.s = value_expr;
    vroot = pv.parse();
    if (!pv.ok) return -1;
  }
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) n *= shape[i];
  if (rank == 0 || n == 0) return 0;
  const Val sv = load_val(dtype, scalar ? scalar : dst, 0);
  int64_t coord[OR_MAXR] = {0};
  for (int64_t idx = 0; idx < n; ++idx) {
    const Val m = eval_lazy(pm, mroot, n_in, ins, coord, rank, flags);
    if (m.i) {
      const Val v = value_expr ? eval_lazy(pv, vroot, n_in, ins, coord, rank, flags) : sv;
      store_val(v, dst, idx);
    }
    for (int i = rank - 1; i >= 0; --i) {
      if (++coord[i] < shape[i]) break;
      coord[i] = 0;
    }
  }
  return 0;
}

// ===========================================================================
// Reductions — Enumerable#sum / #max over `each` (src/multi_indexable.cr:18,
// :656-658; NArray#each src/n_array.cr:556-558) and the slices idiom
// (src/multi_indexable.cr:742-786).
// ===========================================================================

// Enumerable#sum for Float64 elements continued from a running total `start`
// (the sequential sum of a long sequence, fed in pieces).
double or_sum_seq_f64_from(const double *x, int64_t n, double start) {
  double s = start;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  return s;
}

// Enumerable#sum for Float64 elements: sequential lex-order sum from 0.0.
double or_sum_seq_f64(const double *x, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  return s;
}

// Enumerable#sum for Float32 elements: sequential lex-order sum in Float32
// from Float32 zero (the additive identity of the element type).
float or_sum_seq_f32(const float *x, int64_t n) {
  float s = 0.0f;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  return s;
}

// Enumerable#max: keeps the first element, replaces on `>`.
double or_max_seq_f64(const double *x, int64_t n) {
  if (n == 0) return 0.0;
  double m = x[0];
  for (int64_t i = 1; i < n; ++i)
    if (x[i] > m) m = x[i];
  return m;
}

// Enumerable#sum over Int32 with checked `+`: returns OverflowError (-5) if
// any running total leaves the Int32 range.
int or_sum_i32(const int32_t *x, int64_t n, int64_t *out) {
  int64_t s = 0;
  int rc = OK;
  for (int64_t i = 0; i < n; ++i) {
    s += x[i];
    if (!fits32(s)) rc = E_OVERFLOW;
  }
  *out = s;
  return rc;
}

// Correctly rounded sum (Shewchuk's exact partials, the published algorithm
// behind Python's math.fsum, with its half-way rounding correction).  This is
// the FP64 reference a compensated GPU sum is measured against; it is not
// ph-core's (sequential) result, whose own error grows with n.  Non-finite
// terms are summed apart, as math.fsum does: any NaN gives NaN, infinities
// of one sign that infinity, of both signs NaN (where math.fsum raises).
double or_sum_exact(const double *x, int64_t n) {
  std::vector<double> p;
  p.reserve(64);
  double special = 0.0;
  bool nonfinite = false;
  for (int64_t k = 0; k < n; ++k) {
    double v = x[k];
    if (!std::isfinite(v)) {
      special += v;
      nonfinite = true;
      continue;
    }
    size_t i = 0;
    for (size_t j = 0; j < p.size(); ++j) {
      double y = p[j];
      if (std::fabs(v) < std::fabs(y)) { double t = v; v = y; y = t; }
      double hi = v + y;
      double lo = y - (hi - v);
      if (lo != 0.0) p[i++] = lo;
      v = hi;
    }
    p.resize(i);
    p.push_back(v);
  }
  if (nonfinite) return special;
  if (p.empty()) return 0.0;
  size_t m = p.size();
  double hi = p[--m];
  double lo = 0.0;
  while (m > 0) {
    double v = hi;
    double y = p[--m];
    hi = v + y;
    double yr = hi - v;
    lo = y - yr;
    if (lo != 0.0) break;
  }
  if (m > 0 && ((lo < 0.0 && p[m - 1] < 0.0) || (lo > 0.0 && p[m - 1] > 0.0))) {
    double y = lo * 2.0;
    double v = hi + y;
    double yr = v - hi;
    if (y == yr) hi = v;
  }
  return hi;
}

// slices(axis).map(&.sum / &.max) for a row-major array of doubles: result j
// reduces the elements with coordinate `axis` == j, each slice in lex order
// (ChunkIterator -> RegionIterator, src/iterators/region_iterator.cr:56-146,
// one unsafe_fetch_chunk per slice).  mode 0: sequential sum, 1: max,
// 2: correctly rounded sum.
int or_slices_reduce(const double *x, const int64_t *shape, int rank, int axis,
                     int mode, double *out) {
  int64_t outer = 1, inner = 1;
  for (int i = 0; i < axis; ++i) outer *= shape[i];
  for (int i = axis + 1; i < rank; ++i) inner *= shape[i];
  int64_t len = shape[axis];
  std::vector<double> buf((size_t)(outer * inner));
  for (int64_t j = 0; j < len; ++j) {
    size_t k = 0;
    for (int64_t o = 0; o < outer; ++o)
      for (int64_t i = 0; i < inner; ++i)
        buf[k++] = x[(o * len + j) * inner + i];
    if (mode == 0) out[j] = or_sum_seq_f64(buf.data(), (int64_t)buf.size());
    else if (mode == 1) out[j] = or_max_seq_f64(buf.data(), (int64_t)buf.size());
    else out[j] = or_sum_exact(buf.data(), (int64_t)buf.size());
  }
  return OK;
}

// ===========================================================================
// NArray.concatenate_to_slice (src/n_array.cr:727-751): the output has the
// inputs' common shape with `axis` summed; for each of the
// prod(shape[:axis]) chunks, each input in turn contributes its next
// shape_i[axis] * prod(shape[axis+1:]) elements in lexicographic order.
// bufs[i] is input i (row-major, esize-byte elements), shapes is n x rank.
// Returns the number of elements written.
// ===========================================================================
int64_t or_concatenate(int n, const void *const *bufs, const int64_t *shapes, int rank, int esize, int axis,
                       void *out) {
  int64_t inner = 1, chunks = 1;
  for (int d = axis + 1; d < rank; ++d) inner *= shapes[d];
  for (int d = 0; d < axis; ++d) chunks *= shapes[d];
  std::vector<const char *> it((size_t)n);
  for (int i = 0; i < n; ++i) it[(size_t)i] = (const char *)bufs[i];
  char *o = (char *)out;
  int64_t written = 0;
  for (int64_t c = 0; c < chunks; ++c)
    for (int i = 0; i < n; ++i) {
      const int64_t len = shapes[(int64_t)i * rank + axis] * inner;
      std::memcpy(o, it[(size_t)i], (size_t)(len * esize));
      o += len * esize;
      it[(size_t)i] += len * esize;
      written += len;
    }
  return written;
}

// ===========================================================================
// Batched small arrays: the host loop ph-core would run over a list of
// Float32 NArrays (no batched API exists in the reference):
//   y_k = x_k.map { |e| e + s_k }            n_array.cr:589-595 (Float32 +)
//   sums[k] = x_k.sum                        Enumerable#sum, Float32 sequence
//   z_k = x_k.view(0..2.., ...).to_narr      multi_indexable.cr:881-887,
//                                            view.cr:123 (region 0..2.. per axis)
// Arrays are packed: array k at element xoffs[k] of x / y, its slice at
// zoffs[k] of z.  shapes is n x 3, row k holding ranks[k] extents.
// ===========================================================================
int or_batch_loop(int32_t n, const int32_t *ranks, const int64_t *shapes, const int64_t *xoffs,
                  const int64_t *zoffs, const float *scalars, const float *x, float *y, float *z,
                  double *sums, char *msg, int msglen) {
  for (int32_t k = 0; k < n; ++k) {
    const int r = ranks[k];
    const int64_t *shp = shapes + 3 * (int64_t)k;
    int64_t cnt = 1;
    for (int i = 0; i < r; ++i) cnt *= shp[i];
    const float *xk = x + xoffs[k];
    float *yk = y + xoffs[k];
    const float s = scalars[k];
    for (int64_t i = 0; i < cnt; ++i) yk[i] = xk[i] + s;
    sums[k] = (double)or_sum_seq_f32(xk, cnt);
    or_lit lits[OR_MAXR];
    std::memset(lits, 0, sizeof(lits));
    for (int i = 0; i < r; ++i) {
      lits[i].kind = 0;
      lits[i].has_first = 1;
      lits[i].first = 0;
      lits[i].has_step = 1;
      lits[i].step = 2;
    }
    or_region reg;
    int rc = or_region_new(lits, r, shp, r, 1, &reg, msg, msglen);
    if (rc != OK) return rc;
    or_gather(xk, shp, r, 4, &reg, z + zoffs[k]);
  }
  return OK;
}

// ===========================================================================
// Heat equation
// ===========================================================================

// examples/heat_equation.cr, run exactly as written: 21 points (LENGTH //
// SPACING + 1, :12), T_LEFT 0 / T_RIGHT 100 / T_INITIAL 20 (:5-7, :22-24),
// COEFF (:20), steps = (100 / 0.01).to_i32 + 1 = 10001 (:27), update_temp
// (:38-51).  `steps` may be given to run fewer; out receives 21 values.
int or_heat1d_reference(int64_t steps, double *out) {
  const double spacing = 0.05, timestep = 0.01;
  // NUM_POINTS = LENGTH // SPACING + 1: Int32 // Float64 -> Float64 floor
  const int n = (int)(std::floor(1.0 / spacing) + 1.0);
  // COEFF = (237 * 0.01) / (2700 * 900 * (0.05 ** 2)); 0.05 ** 2 is powi(0.05, 2)
  const double coeff = (237 * timestep) / ((double)(2700 * 900) * powi_d(spacing, 2));
  if (steps < 0) steps = (int64_t)(int32_t)(100.0 / timestep) + 1;
  std::vector<double> s((size_t)n, 20.0), d((size_t)n);
  s[0] = 0.0;
  s[(size_t)n - 1] = 100.0;
  for (int64_t t = 0; t < steps; ++t) {
    std::fill(d.begin(), d.end(), 0.0);
    d[0] = (s[1] - s[0]) * coeff;
    d[(size_t)n - 1] = (s[(size_t)n - 2] - s[(size_t)n - 1]) * coeff;
    for (int i = 0; i < n - 2; ++i)
      d[(size_t)i + 1] = ((s[(size_t)i] - 2 * s[(size_t)i + 1]) + s[(size_t)i + 2]) * coeff;
    for (int i = 0; i < n; ++i) s[(size_t)i] = s[(size_t)i] + d[(size_t)i];
  }
  for (int i = 0; i < n; ++i) out[i] = s[(size_t)i];
  return n;
}

// 2-D generalisation used for the GPU stencil: 5-point FTCS, cells outside
// the ny x nx grid read 0 (Dirichlet zero), the per-cell operation order of
// heat_equation.cr:47/:50 extended to two axes:
//   d = (((uN - 4u) + uS) + uW) + uE;   u' = u + d * r
// dtype 2 (F64): exactly that, every operation rounded.  dtype 1 (F32, all
// arithmetic in float, r rounded to float): the build's Float32 variant of
// the workload (the reference's example is Float64 only), whose last two
// operations round once, u' = fma(d, r, u), as the device computes them.
// energy (may be NULL): FP64 sum of the field (correctly rounded) at steps
// 0, every, 2*every, ...
extern "C++" {
template <typename T>
static void heat2d_t(T *u, int64_t ny, int64_t nx, T r, int64_t steps,
                     int64_t every, double *energy) {
  std::vector<T> v((size_t)(ny * nx));
  std::vector<double> tmp;
  int64_t ei = 0;
  auto record = [&]() {
    tmp.resize((size_t)(ny * nx));
    for (int64_t k = 0; k < ny * nx; ++k) tmp[(size_t)k] = (double)u[k];
    energy[ei++] = or_sum_exact(tmp.data(), ny * nx);
  };
  if (energy && every > 0) record();
  for (int64_t t = 0; t < steps; ++t) {
    for (int64_t i = 0; i < ny; ++i) {
      for (int64_t j = 0; j < nx; ++j) {
        T c = u[i * nx + j];
        T n = i > 0 ? u[(i - 1) * nx + j] : T(0);
        T s = i + 1 < ny ? u[(i + 1) * nx + j] : T(0);
        T w = j > 0 ? u[i * nx + j - 1] : T(0);
        T e = j + 1 < nx ? u[i * nx + j + 1] : T(0);
        T d = (((n - T(4) * c) + s) + w) + e;
        if constexpr (sizeof(T) == 4) v[(size_t)(i * nx + j)] = std::fma(d, r, c);
        else v[(size_t)(i * nx + j)] = c + d * r;
      }
    }
    std::memcpy(u, v.data(), sizeof(T) * (size_t)(ny * nx));
    if (energy && every > 0 && (t + 1) % every == 0) record();
  }
}

}  // extern "C++"

int or_heat2d(int dtype, void *u, int64_t ny, int64_t nx, double r,
              int64_t steps, int64_t every, double *energy) {
  if (dtype == DT_F32) heat2d_t<float>((float *)u, ny, nx, (float)r, steps, every, energy);
  else if (dtype == DT_F64) heat2d_t<double>((double *)u, ny, nx, r, steps, every, energy);
  else return E_ARGUMENT;
  return OK;
}

// ===========================================================================
// Seeded inputs: counter-based splitmix64 (Steele, Lea & Flood 2014).  Draw k
// of stream `seed` is the (k+1)-th output of splitmix64 started at `seed`.
// ===========================================================================

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static uint64_t draw(uint64_t seed, uint64_t k) {
  return mix64(seed + (k + 1) * 0x9E3779B97F4A7C15ULL);
}

uint64_t or_draw(uint64_t seed, uint64_t k) { return draw(seed, k); }

// dist 0: uniform [a, b); 1: normal(mean a, stddev b) by Box-Muller;
// 2: Int32 uniform [0, 2^31).  BF16 output rounds the Float32 value to
// nearest-even.
int or_fill_random(void *buf, int dtype, int64_t n, uint64_t seed,
                   uint64_t offset, int dist, double a, double b) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t k = offset + (uint64_t)i;
    double x = 0.0;
    int64_t xi = 0;
    if (dist == 0) {
      uint64_t d = draw(seed, k);
      if (dtype == DT_I32) {
        int64_t span = (int64_t)(b - a);
        xi = (int64_t)a + (int64_t)((d >> 11) % (uint64_t)span);
      } else {
        double u = (double)(d >> 11) * 0x1.0p-53;
        x = a + (b - a) * u;
      }
    } else if (dist == 1) {
      uint64_t kp = k & ~1ULL;  // pair (2m, 2m+1) of the absolute draw index
      double u1 = (double)(draw(seed, kp) >> 11) * 0x1.0p-53;
      double u2 = (double)(draw(seed, kp + 1) >> 11) * 0x1.0p-53;
      double r = std::sqrt(-2.0 * std::log(1.0 - u1));
      double t = two_pi * u2;
      double z = (k & 1) ? r * std::sin(t) : r * std::cos(t);
      x = a + b * z;
    } else {
      xi = (int64_t)(draw(seed, k) >> 33);
    }
    switch (dtype) {
      case DT_I32: ((int32_t *)buf)[i] = (int32_t)xi; break;
      case DT_F32: ((float *)buf)[i] = (float)x; break;
      case DT_F64: ((double *)buf)[i] = x; break;
      case DT_BF16: {
        float f = (float)x;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        uint32_t lsb = (u >> 16) & 1;
        u += 0x7FFF + lsb;
        ((uint16_t *)buf)[i] = (uint16_t)(u >> 16);
        break;
      }
      default: return E_ARGUMENT;
    }
  }
  return OK;
}

}  // extern "C"
