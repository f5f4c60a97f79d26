/*
 * phhip.h — C-ABI of the MI355X (gfx950) backend for ph-core's NArray
 * data-parallel path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, never
 * throws, and returns an int status (PH_OK = 0, negative on error; the text of
 * the last error is in ph_last_error()).  This is exactly the surface a Crystal
 * `lib LibPhHip` block binds: validation (region canonicalisation, shape
 * checks, exception types and messages) stays on the host side, above this
 * boundary, as in ph-core, where it is done before any element is touched.
 *
 * Device memory is owned by the caller (ph_alloc / ph_free).  Nothing here
 * retains a pointer after it returns.  All kernels run on one HIP stream per
 * process (ph_get_stream), at the default priority; calls are asynchronous
 * with respect to the host unless the entry point says otherwise.
 */
#ifndef PHHIP_H
#define PHHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PH_MAX_RANK 8
#define PH_MAX_STAGES 2
#define PH_MAX_INSN 48
#define PH_MAX_INPUTS 8
#define PH_NREG 8

/* Element types.  BF16 is a storage type only: it is widened to Float32 on
 * load (ph-core has no BF16; a BF16 array behaves as Float32 data). BOOL is
 * one byte, 0 or 1. */
typedef enum {
  PH_I32 = 0,
  PH_F32 = 1,
  PH_F64 = 2,
  PH_BF16 = 3,
  PH_BOOL = 4
} ph_dtype;

/* Status codes, one per exception class the reference raises on this path
 * (src/exceptions/exceptions.cr:4-22 and Crystal's IndexError, ArgumentError,
 * OverflowError, DivisionByZeroError). */
enum {
  PH_OK = 0,
  PH_E_INDEX = -1,
  PH_E_SHAPE = -2,
  PH_E_DIMENSION = -3,
  PH_E_ARGUMENT = -4,
  PH_E_OVERFLOW = -5,
  PH_E_DIVZERO = -6,
  PH_E_HIP = -10,
  PH_E_NOINIT = -11,
  PH_E_UNSUPPORTED = -12,
  PH_E_OOM = -13
};

/* Per-launch element flags reported by ph_map / ph_reduce (Crystal raises on
 * these; the host side turns them into the matching exception). */
enum {
  PH_FLAG_OVERFLOW = 1, /* checked Int32 + - * ** or unary - overflowed   */
  PH_FLAG_DIVZERO = 2,  /* Int32 // or % by zero                          */
  PH_FLAG_NEGEXP = 4    /* Int32 ** with a negative exponent (ArgumentError)*/
};

/* ---------------------------------------------------------------------------
 * View descriptor.
 *
 * A logical coordinate c (rank `view.rank`, extents `view.shape`) maps to a
 * buffer element index by
 *     idx = view.offset + sum_i c_i * view.stride[i]
 *     for s in stage[0 .. nstage-1]:
 *         z   = delinearize(idx, s.shape[0..s.rank-1])   (row-major)
 *         idx = s.offset + sum_j z_j * s.stride[j]
 * Strides are signed element strides; 0 repeats an element (broadcast).
 * Region, Permute and Reverse transforms (their apply methods at
 * src/view_util/transforms.cr:215-221, :267-270 and :298-304) fold into the
 * affine map; Reshape (:184-188) and tile (src/multi_indexable.cr:818-827)
 * need a stage.  `extent` is the element
 * count of the underlying buffer; the library rejects a descriptor that can
 * address outside [0, extent).
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t rank;
  int32_t pad_;
  int64_t offset;
  int64_t shape[PH_MAX_RANK];
  int64_t stride[PH_MAX_RANK];
} ph_affine;

typedef struct {
  int32_t dtype;
  int32_t nstage;
  int64_t extent;
  ph_affine view;
  ph_affine stage[PH_MAX_STAGES];
} ph_view_desc;

/* ---------------------------------------------------------------------------
 * Element program: the compiled form of an elementwise expression (the
 * operator macros of src/multi_indexable.cr:931-985, scalar operands
 * src/multi_indexable.cr:947-951 and src/patches/number.cr:5-15, and a
 * ProcView block src/proc_view.cr:28-30).  Registers hold one value each
 * (Int32, Float32, Float64 or Bool, tagged by the instruction's dtype);
 * results round to the instruction's dtype exactly as Crystal does, with no
 * contraction into fused multiply-adds.
 * ------------------------------------------------------------------------- */
typedef enum {
  PH_OP_LOAD = 0,  /* r[dst] = input[a] widened to dt (BF16 -> F32)       */
  PH_OP_CONST,     /* r[dst] = imm (dt)                                    */
  PH_OP_CAST,      /* r[dst] = (dt) r[a], source dtype in c                */
  PH_OP_ADD,       /* dt I32: checked; F32/F64: IEEE                        */
  PH_OP_SUB,
  PH_OP_MUL,
  PH_OP_WADD,      /* Int32 &+ &- &* (wrapping)                             */
  PH_OP_WSUB,
  PH_OP_WMUL,
  PH_OP_DIV,       /* F32/F64 true division                                 */
  PH_OP_FLOORDIV,  /* // : I32 floored (div by 0 flagged), F floor(a/b)     */
  PH_OP_MOD,       /* %  : I32 floored modulo, F a - b*floor(a/b)           */
  PH_OP_POWI,      /* F ** Int32 by binary powering (llvm.powi)             */
  PH_OP_IPOW,      /* Int32 ** Int32, checked                               */
  PH_OP_WIPOW,     /* Int32 &** Int32, wrapping                             */
  PH_OP_POW,       /* F ** F (libm pow)                                     */
  PH_OP_NEG,       /* unary - (checked for I32)                             */
  PH_OP_BNOT,      /* unary ~ (I32)                                         */
  PH_OP_AND,       /* & | ^ (I32, BOOL)                                     */
  PH_OP_OR,
  PH_OP_XOR,
  PH_OP_GT,        /* comparisons of two dt operands -> BOOL                */
  PH_OP_LT,
  PH_OP_GE,
  PH_OP_LE,
  PH_OP_EQ,
  PH_OP_NE,
  PH_OP_CMP,       /* <=> -> I32 (-1, 0, 1)                                 */
  PH_OP_SQRT,      /* F sqrt (correctly rounded)                            */
  PH_OP_ABS,
  PH_OP_SELECT,    /* r[dst] = r[c] ? r[a] : r[b]                           */
  PH_OP_INDEX,     /* r[dst] = the element's lexicographic index (Int32);   */
                   /* ph_map only (map_with_index, src/n_array.cr:605-611) */
  PH_OP__COUNT
} ph_opcode;

typedef struct {
  uint8_t op;
  uint8_t dt;
  uint8_t dst;
  uint8_t a;
  uint8_t b;
  uint8_t c;
  uint16_t pad_;
  union {
    int64_t i;
    double f;
  } imm;
} ph_insn;

typedef struct {
  int32_t n_insn;
  int32_t n_in;
  int32_t out_reg;
  int32_t out_dtype;
  ph_insn insn[PH_MAX_INSN];
} ph_prog;

/* ---------------------------------------------------------------------------
 * Runtime
 * ------------------------------------------------------------------------- */

/* Select `device` and create the library's stream.  Replaces nothing in the
 * reference (ph-core is CPU-only, src/n_array.cr:20 owns a host Slice). */
int ph_init(int device);
int ph_device_count(int *n);
const char *ph_last_error(void);

/* Device buffers backing an NArray (src/n_array.cr:20 `@buffer : Slice(T)`). */
int ph_alloc(void **ptr, size_t bytes);
int ph_free(void *ptr);
/* Synchronous copies (the stream is drained first). */
int ph_h2d(void *dst, const void *src, size_t bytes);
int ph_d2h(void *dst, const void *src, size_t bytes);
/* Asynchronous device-to-device copy on the library stream. */
int ph_d2d(void *dst, const void *src, size_t bytes);
int ph_sync(void);
/* The hipStream_t every ph_* kernel is launched on. */
int ph_get_stream(void **stream);
/* HIP events on the library stream, for timing. */
int ph_event_create(void **ev);
int ph_event_record(void *ev);
int ph_event_elapsed(float *ms, void *start, void *stop);
int ph_event_destroy(void *ev);

/* ---------------------------------------------------------------------------
 * Data generation.  Counter-based splitmix64 stream: draw k of stream `seed`
 * is splitmix64 output number k+1 from state seed; element i takes draw
 * offset+i (normal: draws offset+2*(i/2), +1 feed one Box-Muller pair).
 * dist: 0 uniform [a,b) (F: top 53 bits; I32: a + (draw>>11) mod (b-a)),
 *       1 normal(a = mean, b = stddev), 2 Int32 uniform [0, 2^31) (draw>>33).
 * No reference counterpart (ph-core's specs use literals); inputs of the
 * benchmark workloads are generated on the device with it.
 * ------------------------------------------------------------------------- */
int ph_fill_random(void *buf, int32_t dtype, int64_t n, uint64_t seed,
                   uint64_t offset, int32_t dist, double a, double b);

/* ---------------------------------------------------------------------------
 * Fill every element addressed by `dst` with one scalar (`value` points to one
 * element of dst->dtype).
 * Replaces NArray.fill (src/n_array.cr:230-232), NArray#unsafe_set_chunk(
 * region, value) (src/n_array.cr:495-500) and MultiWritable#unsafe_set_chunk(
 * region, value) (src/multi_writable.cr:40-44).
 * ------------------------------------------------------------------------- */
int ph_fill(void *buf, const ph_view_desc *dst, const void *value);

/* ---------------------------------------------------------------------------
 * Strided copy over a common logical shape: for every coordinate k,
 *     dst[dd(k)] = src[sd(k)]      (same dtype; dd and sd have equal shapes)
 * Gather: sd is a view, dd contiguous  — NArray#unsafe_fetch_chunk
 *   (src/n_array.cr:450-453, odometer src/buffered/indexed/lex_iterator.cr:
 *   7-21), View#to_narr (src/view.cr:123-126), MultiIndexable#reshape/
 *   permute/reverse (src/multi_indexable.cr:795-803).
 * Scatter: dd is a view, sd contiguous — NArray#unsafe_set_chunk(region,
 *   src) (src/n_array.cr:484-492), MultiWritable#unsafe_set_chunk
 *   (src/multi_writable.cr:29-36) through MutableView#unsafe_set_element
 *   (src/mutable_view.cr:16-18).
 * A source that shares the destination's buffer (src == dst) is read as
 * the reference's element loop reads it: element k (lexicographic over the
 * shape) sees the writes of the elements before it, since the loop pulls
 * `src.each` one element at a time (src/n_array.cr:484-492).  The library
 * finds whether any element reads another's destination and runs the
 * parallel copy when none does; otherwise it resolves the order exactly
 * (pointer jumping along the chains of reads after writes).  A source that
 * overlaps the destination through another base pointer is PH_E_ARGUMENT.
 * flags: must be 0.
 * ------------------------------------------------------------------------- */
int ph_strided_copy(void *dst, const ph_view_desc *dd, const void *src,
                    const ph_view_desc *sd, int32_t flags);

/* ---------------------------------------------------------------------------
 * Fused fill + scatter over a whole row-major destination buffer of shape
 * dst_shape[rank]: element b gets src[sd(c)] if, on every axis a,
 *     b_a = first[a] + c_{axis[a]} * step[a],  0 <= c_{axis[a]} < count[a]
 * (axis[a] = -1: the axis is fixed at first[a]); every other element gets
 * *fill.  One pass writes the destination once and reads the source once.
 * Replaces NArray.fill(shape, value) (src/n_array.cr:230-232) followed by a
 * scatter through a region / MutableView (src/multi_writable.cr:55-63,
 * src/mutable_view.cr:16-18) into that fresh array.
 * ------------------------------------------------------------------------- */
int ph_fill_scatter(void *dst, int32_t dtype, const int64_t *dst_shape,
                    int32_t rank, const int64_t *first, const int64_t *step,
                    const int64_t *count, const int32_t *axis,
                    const void *src, const ph_view_desc *sd,
                    const void *fill);

/* ---------------------------------------------------------------------------
 * Elementwise map: out[k] = prog(in_0[sd_0(k)], ..., in_{n-1}[sd_{n-1}(k)])
 * for every coordinate k of out_desc's logical shape; every in_desc has that
 * shape (broadcast is expressed by stride 0 / tile stages).  `flags_out`
 * (host pointer, may be NULL) receives the OR of PH_FLAG_* raised by any
 * element; the call then synchronises.
 * Replaces MultiIndexable#map_with / each_with (src/multi_indexable.cr:
 * 1026-1102), the operator macros (:931-985), NArray#map / map! (src/
 * n_array.cr:589-628), scalar-first patches (src/patches/number.cr:5-21),
 * and ProcView#to_narr (src/proc_view.cr:28-30 + src/view.cr:123-126).
 * In-place map (map!) passes out == in_0 with equal contiguous descriptors.
 * Inputs may read the output's buffer (the same base pointer, the output's
 * element type) through any descriptor: element k then sees the writes of
 * the elements before it in lexicographic order, as the reference's loops
 * do when a source or mask is a view of the array being written
 * (src/n_array.cr:484-492, :523-536).  When no element reads another's
 * output the map runs in one parallel pass; otherwise the order is resolved
 * by repeated passes to a fixed point (at most 12; 3 when only one input
 * reads the buffer).  With one input on the buffer, deeper dependencies
 * run a device thread per chain when no element has two dependents (with
 * no passes at all when that input's affine view reads each address at
 * most once); else, for the tree of dependents, a thread per chain of the
 * elements that others read (when those form chains) and one pass for the
 * rest, or a pass per depth level (when the levels average 16 elements or
 * more and number at most 65536).  Other deep maps run a sequential loop
 * on one device thread.  An input that overlaps the output through another
 * base pointer is PH_E_ARGUMENT.
 * A program with PH_OP_INDEX (NArray#map_with_index / map_with_index! /
 * map_with_coord, src/n_array.cr:605-650) sees each element's lexicographic
 * position in out_desc's shape as an Int32; a shape of more than 2^31 - 1
 * elements is PH_E_OVERFLOW, since Crystal's Int32 index cannot reach it.
 * ------------------------------------------------------------------------- */
int ph_map(const ph_prog *prog, int32_t n_in, const void *const *ins,
           const ph_view_desc *in_descs, void *out,
           const ph_view_desc *out_desc, int32_t *flags_out);

/* ---------------------------------------------------------------------------
 * Buffer element index of every coordinate of a view, lexicographic:
 *     out[k] = d(k)      (device int64 array of desc_count(d) entries)
 * ComposedTransform#apply (src/view_util/transforms.cr:28-101) for the part
 * of a view chain that one descriptor folds.  A chain longer than a
 * descriptor holds composes these arrays: the index array of its first
 * descriptor, gathered (ph_strided_copy of 8-byte elements) through the
 * descriptors of the rest, is the chain's own index array.
 * ------------------------------------------------------------------------- */
int ph_view_index(int64_t *out, const ph_view_desc *d);

/* ---------------------------------------------------------------------------
 * ph_map whose operands may be addressed through index arrays (views whose
 * transform chain is longer than one descriptor holds, src/view_util/
 * transforms.cr:28-101).  For element k (lexicographic over out_desc's
 * shape):
 *   input j reads  ins[j][in_index[j][k]]  when in_index && in_index[j],
 *                  ins[j][in_descs[j](k)]  otherwise;
 *   the result goes to out[out_index[k]] when out_index, else
 *   out[out_desc(k)].
 * Index arrays are device int64 arrays of desc_count(out_desc) entries.  An
 * indexed operand's descriptor gives its dtype, buffer extent and logical
 * shape; its offset, strides and stages are ignored.  prog == NULL with
 * n_in == 1 copies input 0 (out's dtype or one of its size, bits unchanged).
 * Every index must lie in [0, extent) (PH_E_INDEX, nothing written) and
 * out_index must name each element at most once (PH_E_ARGUMENT, nothing
 * written; a MutableView chain is one-to-one).  Inputs on the output's
 * buffer are read in the reference's element order, as in ph_map: element
 * k sees the writes of the elements before it.
 * Replaces NArray#unsafe_set_chunk(region, src) (src/n_array.cr:484-492)
 * and NArray#[]=(mask, value) (:523-536) where the source, the mask or the
 * destination (a MutableView, src/mutable_view.cr:16-18) is such a view of
 * the array being written.  Scratch: 8 bytes per element of out's buffer
 * when out_index is given (the inverse of the write map).
 * ------------------------------------------------------------------------- */
int ph_map_indexed(const ph_prog *prog, int32_t n_in, const void *const *ins,
                   const ph_view_desc *in_descs, const int64_t *const *in_index,
                   void *out, const ph_view_desc *out_desc,
                   const int64_t *out_index, int32_t *flags_out);

/* ---------------------------------------------------------------------------
 * Fused map + reduction of prog's value over a view of rank >= 1.
 * keep_axis = -1: one result (Enumerable#sum / #max over `each`,
 *   src/multi_indexable.cr:18, :656-658).
 * keep_axis = k : shape[k] results, result j reducing the slice with
 *   coordinate k fixed at j (`slices(k).map(&.sum)`, src/multi_indexable.cr:
 *   742-786).
 * ops: bit 0 sum, bit 1 max.  acc: accumulation mode for float values.
 *   Each tile of at most 1024 rows sums its values in FP64; the tiles'
 *   partial sums are carried as compensated pairs and combined with
 *   error-free TwoSum, pairs of floats under PH_ACC_F32C and of doubles
 *   under PH_ACC_F64.  Int32/Bool values are summed exactly in 64 bits; an Int32
 *   sum whose sequential (lexicographic) running total ever leaves the Int32
 *   range sets PH_FLAG_OVERFLOW, as Crystal's checked `+` would raise.
 * out_sum / out_max: device arrays of doubles (count results each; may be
 *   NULL when the op is not requested).
 * ------------------------------------------------------------------------- */
#define PH_RED_SUM 1
#define PH_RED_MAX 2
#define PH_ACC_F32C 0
#define PH_ACC_F64 1
int ph_reduce(const ph_prog *prog, int32_t n_in, const void *const *ins,
              const ph_view_desc *in_descs, int32_t keep_axis, int32_t ops,
              int32_t acc, double *out_sum, double *out_max,
              int32_t *flags_out);

/* One-pass 2-D variant: for a rows x cols view, writes in one read of the
 * data full sum/max (1 each), per-row sum/max (slices(0), `rows` each) and
 * per-column sum/max (slices(1), `cols` each).  Any output may be NULL. */
int ph_reduce2d_all(const ph_prog *prog, int32_t n_in, const void *const *ins,
                    const ph_view_desc *in_descs, int32_t acc,
                    double *full_sum, double *full_max, double *row_sum,
                    double *row_max, double *col_sum, double *col_max,
                    int32_t *flags_out);

/* ---------------------------------------------------------------------------
 * Explicit 2-D heat equation, 5-point FTCS stencil, Dirichlet-zero outside
 * the ny x nx grid (cells beyond the edge read 0):
 *   u'[i,j] = u[i,j] + ((((uN - 4u) + uS) + uW) + uE) * r
 * the 2-D generalisation of examples/heat_equation.cr:38-51 (same operation
 * order as :47, `(s[i-1] - 2*s[i] + s[i+1]) * COEFF`, then `el + diff` :50).
 * dtype PH_F64 rounds every operation as written (the reference's example
 * is Float64).  dtype PH_F32, the build's own Float32 variant, computes the
 * last multiply and add with one rounding, u' = fma(d, r, u); its accuracy
 * is stated against the Float64 run.  u holds the initial field and
 * receives the final one; tmp is scratch of the same size.  Up to eight
 * (Float64) or twelve (Float32) steps run per pass over the grid (temporal
 * blocking), bit-identical to single steps.  If
 * energy_every > 0, energy_out
 * (device, steps/energy_every + 1 doubles) receives the FP64 sum of u at
 * steps 0, energy_every, 2*energy_every, ...
 * ------------------------------------------------------------------------- */
int ph_heat2d(int32_t dtype, void *u, void *tmp, int64_t ny, int64_t nx,
              double r, int64_t steps, int64_t energy_every,
              double *energy_out);

/* ph_heat2d without any allocation.  It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and the same results, bit for bit.  ph_heat2d keeps the partial
 * sums of its FP64 energy reductions in library scratch that grows on
 * demand; this variant keeps them in a fixed 32 KiB device array (2048
 * groups, the most the reduction uses for any grid), whose address ph_init
 * looks up on its device, and it skips ph_heat2d's one read-back of max|u|
 * (the fused first operation it enables changes no result).  Every call on a
 * device shares that array, and each reduction writes every slot it later
 * reads, so nothing needs clearing; calls that overlap on one device (from
 * several threads or processes sharing the library's state) race. */
int ph_heat2d_noalloc(int32_t dtype, void *u, void *tmp, int64_t ny, int64_t nx,
                      double r, int64_t steps, int64_t energy_every,
                      double *energy_out);

/* Sum of squares of (a - b) and of b, both in FP64 (a, b of dtype F32 or
 * F64, n elements): the relative L2 distance report of a reduced-precision
 * field against an FP64 one.  out: device, 2 doubles. */
int ph_l2_diff(int32_t dtype_a, const void *a, int32_t dtype_b, const void *b,
               int64_t n, double *out);

/* ---------------------------------------------------------------------------
 * Batched small arrays.  No reference counterpart: ph-core loops over its
 * arrays on the host.  For each Float32 array x_k (rank 1..3) of a batch it
 * computes, in one persistent-kernel pass over the packed buffer,
 *   y_k = x_k + s_k                      per element, like `x_k.map &.+(s_k)`
 *                                        (n_array.cr:589-595) in Float32;
 *   sums[k] = x_k.sum                    Enumerable#sum over `each` (FP32
 *                                        Neumaier accumulation, returned as
 *                                        FP64);
 *   z_k = x_k.view(...).to_narr          every axis from 0 with step 2
 *                                        (multi_indexable.cr:881-887,
 *                                        view.cr:123).
 * The caller packs x_k row-major at the offsets ph_batch_sizes reports
 * (16-byte aligned), and receives y at the same offsets and z_k packed
 * contiguously.  Sums are deterministic (each array's partials are folded in
 * a fixed order whatever the schedule).
 * ------------------------------------------------------------------------- */
/* ranks[n], shapes[3*n] (row k: its first ranks[k] entries, each in
 * [1, 2^30], at most 2^31-1 elements per array), scalars[n].  An extent of 0
 * is PH_E_INDEX: the step-2 view of an empty axis raises IndexError in the
 * reference (range_syntax.cr:120-122).  x, y, z, sums
 * are device pointers (may be NULL here and bound later with ph_batch_bind).
 * The batch owns its descriptor table and work queue on the device. */
int ph_batch_create(void **handle, int32_t n, const int32_t *ranks,
                    const int64_t *shapes, const float *x, const float *scalars,
                    float *y, float *z, double *sums);
/* Element counts of the packed x/y and z buffers, and per-array offsets
 * (offsets_x / offsets_z may be NULL, else n entries each).  x and y must
 * hold xy_elements (every array padded to a multiple of 4 elements; the
 * kernel reads and writes whole 16-byte vectors) and be 16-byte aligned. */
int ph_batch_sizes(void *handle, int64_t *xy_elements, int64_t *z_elements,
                   int64_t *offsets_x, int64_t *offsets_z);
int ph_batch_bind(void *handle, const float *x, float *y, float *z,
                  double *sums);
/* mode 0: one replay of the HIP graph holding the whole batch (captured on
 * first use); 1: the same three kernels launched directly; 2: the unbatched
 * form, three kernels per array (add, sum, slice copy). */
int ph_batch_run(void *handle, int32_t mode);
/* Times candidate schedules on the bound buffers: work-item size (an array
 * larger than that is split into several items, each a wavefront's work),
 * workgroups per CU, and item assignment (a device queue, or a static
 * largest-first round robin); keeps the fastest, cached per histogram of
 * array sizes, and reports it (any output may be NULL).  Overwrites y, z and
 * sums. */
int ph_batch_tune(void *handle, int64_t *chunk_out, int32_t *grid_out,
                  int32_t *dynamic_out);
/* Sets a schedule explicitly (chunk a multiple of 64 elements, grid
 * workgroups of 256 threads, dynamic 1 = queue, 0 = static). */
int ph_batch_schedule(void *handle, int64_t chunk, int32_t grid,
                      int32_t dynamic);
int ph_batch_destroy(void *handle);

/* ---------------------------------------------------------------------------
 * General batches.  No reference counterpart: ph-core runs per-array work in
 * a host loop over its NArrays.  Record k describes one array x_k (dtype
 * x_dtype: I32, F32 or F64; rank 1..3; extents in [1, 2^30]; fewer than 2^31
 * elements; row-major) and the work done on it:
 *   map >= 0    y_k = progs[map] applied to every element, as
 *               `x_k.map { |v| ... }` (src/n_array.cr:589-595); y has y_dtype
 *               (I32, F32, F64 or Bool), the programs' result type;
 *   red >= 0    v = progs[red] applied to every element (I32, F32 or F64);
 *               with red_ops bit PH_RED_SUM sums[k] = v.sum, with PH_RED_MAX
 *               maxes[k] = v.max (Enumerable#sum / #max over `each`,
 *               src/multi_indexable.cr:18, :656-658); an Int32 sum is exact
 *               and checked like Crystal's `+`: a running total outside
 *               Int32 raises PH_FLAG_OVERFLOW;
 *   slice != 0  z_k = x_k.view(region).to_narr, the region given per axis as
 *               first, step (non-zero; a negative step runs the axis
 *               backwards) and count (src/multi_indexable.cr:881-887,
 *               src/view.cr:123), every index inside the array, else
 *               PH_E_INDEX.
 * A program reads input 0, the element (x_dtype), and input 1, the record's
 * scalar converted to x_dtype.  The whole batch runs as one pass over the
 * packed x: one persistent kernel that hipRTC generates for the batch's
 * programs, fed from a device table of work items (largest first), plus a
 * small kernel that folds each array's partial results; mode 0 replays them
 * as a HIP graph.  Sums are FP64 per lane and folded in a fixed order, so
 * they do not depend on the schedule; maxima are exactly the `>` loop's
 * result (the first of equal maxima, so the first zero's sign; NaN when the
 * first value is NaN, other NaNs skipped).
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t rank;
  int32_t map;      /* program index for y_k, or -1 */
  int32_t red;      /* program index for the reductions, or -1 */
  int32_t red_ops;  /* PH_RED_SUM | PH_RED_MAX, 0 when red < 0 */
  int32_t slice;    /* 1: copy the region into z_k */
  int32_t pad_;
  int64_t shape[3]; /* the first `rank` entries */
  int64_t first[3], step[3], count[3]; /* the region, `rank` entries */
  double scalar;
} ph_batch_rec;

/* n_progs in [0, 16]; the records may use at most 16 distinct
 * (map, red, red_ops, slice) combinations. */
int ph_gbatch_create(void **handle, int32_t x_dtype, int32_t y_dtype,
                     int32_t n_progs, const ph_prog *progs, int32_t n,
                     const ph_batch_rec *recs);
/* Elements the caller allocates: x and y hold x_elements of their dtypes
 * (each array starts at a multiple of 16 elements, and the kernel reads and
 * writes up to the end of that padding), z holds z_elements of x_dtype.
 * x_k and y_k are at offsets_xy[k], z_k at offsets_z[k] (either may be
 * NULL, else n entries). */
int ph_gbatch_sizes(void *handle, int64_t *x_elements, int64_t *z_elements,
                    int64_t *offsets_xy, int64_t *offsets_z);
/* Runs the batch over device buffers: x, y, z, sums[n], maxes[n] (those the
 * records write nothing into may be NULL; x and y 16-byte aligned).  mode 0:
 * replay of the HIP graph captured for these buffers (captured on first use
 * and again when the buffers change); 1: the kernels launched directly; 2:
 * the kernel launched once per array (the work without the batching); 3:
 * each record through ph_map, ph_reduce and ph_strided_copy in turn (those
 * compile a kernel per distinct shape: a cross-check for small batches).
 * flags_out (may be NULL): the PH_FLAG_* bits the
 * programs raised (Int32 overflow, division by zero); y then holds the
 * wrapped values, as ph_map leaves them. */
int ph_gbatch_run(void *handle, const void *x, void *y, void *z, double *sums,
                  double *maxes, int32_t mode, int32_t *flags_out);
/* Times candidate schedules (work-item size, workgroups per CU, queue or
 * static assignment) on the given buffers, keeps the fastest and reports it
 * (outputs may be NULL).  Overwrites y, z, sums and maxes. */
int ph_gbatch_tune(void *handle, const void *x, void *y, void *z,
                   double *sums, double *maxes, int64_t *chunk_out,
                   int32_t *grid_out, int32_t *dynamic_out);
int ph_gbatch_destroy(void *handle);

/* ---------------------------------------------------------------------------
 * Diagnostics (no reference counterpart).
 * ph_compile_check generates and compiles, without launching, the fused
 * kernels ph_map (kind 0) and ph_reduce2d_all (kind 1) would use for `prog`
 * over contiguous inputs of the given dtypes and a rows x cols shape, and
 * (kind 2) the tiled map ph_map uses when input 0 is the transposed view of
 * a cols x rows array (rows and cols of 16 or more).  Kind 3 compiles the
 * kernels of a map whose input 0 reads the output's own buffer transposed
 * (rows == cols; input 0 has the output's dtype), kind 4 those of a copy
 * of in_dtypes[0] elements that reverses a buffer onto itself, kind 5 the
 * general batch's kernels with `prog` as map and (for an I32, F32 or F64
 * result) reduction program (two inputs of one dtype: the element and the scalar;
 * rows and cols unused), kind 6 those of ph_map_indexed with input 0 (of
 * the output's dtype) read from the output's buffer through an index array
 * and the output written through one.  It needs no GPU, so program
 * lowering can be checked on a build host.
 * ------------------------------------------------------------------------- */
int ph_compile_check(const ph_prog *prog, int32_t n_in, const int32_t *in_dtypes,
                     int64_t rows, int64_t cols, int32_t kind);
/* Number of fused-kernel modules compiled (not loaded from cache) so far. */
long ph_jit_compiles(void);
/* Number of device operations issued so far: each kernel launch, or each
 * fixed sequence of launches one entry point issues as a unit (a reduction's
 * tile and finish kernels, a batch pass), counts once.  PH_HIP_FAIL_AFTER=n
 * (read at the first launch) refuses every operation after the first n. */
long ph_launch_count(void);

#ifdef __cplusplus
}
#endif

#endif /* PHHIP_H */
