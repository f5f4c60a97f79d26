// Host side of the MI355X backend: a C++ mirror of ph-core's public API for
// the NArray data-parallel path, on top of the C-ABI in include/phhip.h.
//
// Names, argument meaning and exception behaviour follow the reference:
//   RangeSyntax            src/range_syntax/range_syntax.cr
//   CoordUtil / ShapeUtil  src/coord_util.cr, src/shape_util.cr
//   IndexRegion            src/index_region.cr
//   NArray                 src/n_array.cr + src/multi_indexable.cr +
//                          src/multi_writable.cr (device-resident buffer)
//   View / MutableView /   src/view.cr, src/mutable_view.cr, src/proc_view.cr,
//   ProcView               src/view_util/transforms.cr
// Validation (canonicalisation, shape checks) happens here, before any call
// into the device library, so exception types and messages are decided on
// the host exactly where ph-core decides them.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/phhip.h"

namespace phase {

using Shape = std::vector<int64_t>;
using Coord = std::vector<int64_t>;

// ---------------------------------------------------------------------------
// Exceptions — src/exceptions/exceptions.cr:4-22 and Crystal's stdlib ones.
// ---------------------------------------------------------------------------
struct PhaseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct IndexError : PhaseError {
  using PhaseError::PhaseError;
};
struct ShapeError : PhaseError {
  using PhaseError::PhaseError;
};
struct DimensionError : ShapeError {
  using ShapeError::ShapeError;
};
struct ArgumentError : PhaseError {
  using PhaseError::PhaseError;
};
struct OverflowError : PhaseError {
  using PhaseError::PhaseError;
};
struct DivisionByZeroError : PhaseError {
  using PhaseError::PhaseError;
};
struct EmptyError : PhaseError {  // Enumerable::EmptyError
  using PhaseError::PhaseError;
};
struct DeviceError : PhaseError {  // a failure inside the HIP library
  using PhaseError::PhaseError;
};
struct ParseError : PhaseError {  // JSON::Error / YAML::Error
  using PhaseError::PhaseError;
};

// Raises the exception matching a C-ABI status code.
[[noreturn]] void raise_status(int code, const std::string &context);
void check(int code, const char *context);

enum class DType : int32_t { I32 = PH_I32, F32 = PH_F32, F64 = PH_F64, BF16 = PH_BF16, Bool = PH_BOOL };
const char *dtype_name(DType d);
size_t dtype_size(DType d);
DType compute_dtype(DType storage);  // BF16 computes as Float32

// A scalar operand: Crystal literal types (Int32, Float32, Float64, Bool).
struct Scalar {
  DType dt = DType::I32;
  int64_t i = 0;
  double f = 0.0;
  static Scalar i32(int64_t v);
  static Scalar f32(double v);
  static Scalar f64(double v);
  static Scalar boolean(bool v);
  std::string to_s() const;
};

// ---------------------------------------------------------------------------
// Range literals — the forms parse_range accepts (range_syntax.cr:41-69):
// an Int index, `a..c`, `a...c`, `a..b..c` (= `(a..b)..c`, step b),
// `a..(b..c)`, and `a.step(by: b, to: c, exclusive: e)`.  Written as Crystal
// source text, e.g. "1..3..-1", "..-2..", "(5..-3)..1", "...", "-2".
// ---------------------------------------------------------------------------
struct RangeLiteral {
  bool is_index = false;
  int64_t index = 0;
  std::optional<int64_t> first, last, step;
  bool exclusive = false;
  std::string text;

  static RangeLiteral parse(const std::string &text);
  static RangeLiteral of_index(int64_t i);
  static RangeLiteral full();  // `..`
  std::string to_s() const { return text; }
};
using RegionLiteral = std::vector<RangeLiteral>;

struct CanonicalRange {
  int64_t first, step, last, size;
};

namespace RangeSyntax {
int64_t get_size(int64_t first, int64_t last, int64_t step);
CanonicalRange infer_range(const RangeLiteral &lit, int64_t bound);
CanonicalRange canonicalize_range(const RangeLiteral &lit, int64_t bound);
bool bounded(const RangeLiteral &lit);
void ensure_nonnegative(const RangeLiteral &lit);
}  // namespace RangeSyntax

namespace CoordUtil {
bool has_index(int64_t index, int64_t size);
int64_t canonicalize_index(int64_t index, int64_t size);
Coord canonicalize_coord(const Coord &coord, const Shape &shape);
}  // namespace CoordUtil

namespace ShapeUtil {
bool compatible_shapes(const Shape &a, const Shape &b);
int64_t shape_to_size(const Shape &shape);
}  // namespace ShapeUtil

std::string shape_to_s(const Shape &s);

// ---------------------------------------------------------------------------
// IndexRegion — src/index_region.cr
// ---------------------------------------------------------------------------
class IndexRegion {
 public:
  std::vector<int64_t> first, step, last;
  Shape proper_shape, reduced_shape;
  std::vector<bool> degeneracy;
  bool drop = true;

  // IndexRegion.new(region_literal, bound_shape, drop) (:192-224)
  static IndexRegion from_literal(const RegionLiteral &lit, const Shape &bound, bool drop = true);
  // IndexRegion.new(region_literal, drop) — absolute, no bound (:248-275)
  static IndexRegion absolute(const RegionLiteral &lit, bool drop = true);
  // IndexRegion.new(region_literal, bound_shape?, drop, trim_to:) (:133-168)
  static IndexRegion trimmed(const RegionLiteral &lit, const Shape *bound, bool drop, const Shape &trim_to);
  // IndexRegion.cover (:232-238)
  static IndexRegion cover(const Shape &shape, bool drop = true, const std::vector<bool> *degeneracy = nullptr);
  // IndexRegion.new(region, bound_shape) copy that must fit (:109-115)
  static IndexRegion fitted_copy(const IndexRegion &r, const Shape &bound);

  const Shape &shape() const { return reduced_shape; }
  int dimensions() const { return (int)reduced_shape.size(); }
  int proper_dimensions() const { return (int)proper_shape.size(); }
  int64_t size() const { return ShapeUtil::shape_to_size(reduced_shape); }

  IndexRegion &trim(const Shape &bound);           // trim! (:502-515)
  IndexRegion &reverse();                          // reverse! (:533-537)
  IndexRegion &translate(const std::vector<int64_t> &offset);  // translate! (:577-585)
  IndexRegion &unsafe_translate(const std::vector<int64_t> &offset);
  bool includes(const Coord &coord) const;         // includes? (:443-456)
  bool fits_in(const Shape &bound) const;          // fits_in? (:468-478)
  Coord local_to_absolute_unsafe(const Coord &local) const;  // (:621-637)
  Coord local_to_absolute(const Coord &local) const;
  Coord absolute_to_local_unsafe(const Coord &abs) const;    // (:654-664)
  Coord absolute_to_local(const Coord &abs) const;
  std::string to_s() const;                        // (:593-604)
  bool operator==(const IndexRegion &o) const;

 private:
  void compute_reduced();
};

// ---------------------------------------------------------------------------
// Element expressions: the compiled form of a ProcView block or of an
// elementwise operator, with Crystal's numeric semantics (type promotion,
// checked Int32 arithmetic, floored // and %, powi for Float ** Int32).
// ---------------------------------------------------------------------------
struct ExprNode;
class Expr {
 public:
  std::shared_ptr<const ExprNode> node;

  static Expr input(int k, DType storage);  // element of operand k
  static Expr constant(const Scalar &s);
  // The element's lexicographic index (Int32), as Array#map_with_index yields
  // it; only elementwise maps define it.
  static Expr index();
  DType dtype() const;
  // Crystal operator names: "+", "-", "*", "/", "//", "%", "**", "&+", "&-",
  // "&*", "&**", "&", "|", "^", ">", "<", ">=", "<=", "==", "!=", "<=>"
  static Expr binary(const std::string &op, const Expr &a, const Expr &b);
  // "-", "+", "~", "sqrt", "abs", "to_f32", "to_f64", "to_i32"
  static Expr unary(const std::string &op, const Expr &a);
  static Expr select(const Expr &cond, const Expr &a, const Expr &b);
  // Replace input k by exprs[k] (composition of blocks / operand fusion).
  Expr substitute(const std::vector<Expr> &exprs) const;
  // Index of the highest input referenced + 1.
  int arity() const;
  std::string to_s() const;
  // Lower to a ph_prog over `n_in` inputs (storage dtypes given).
  ph_prog compile(int n_in) const;
};

// ---------------------------------------------------------------------------
// Device storage shared by an NArray and the views over it.  A fill can be
// left pending (NArray.fill): the first scatter into a pending-fill buffer
// writes the fill and the scattered elements in one pass.
// ---------------------------------------------------------------------------
// Device storage of an NArray, with a lazily filled host mirror.  The mirror
// is kept in pages: reading an element on the host fetches only its page (one
// device-to-host copy, then host reads of that page cost nothing), element
// writes land in the mirror and mark their page dirty, and dirty pages are
// written back (one host-to-device copy per run of dirty pages) before the
// device next uses the storage.  A device use that may write the storage
// drops the mirror.  A pending fill is applied to host pages without any
// copy, and issued on the device at its first use there.
struct PendingMap;

struct Storage : std::enable_shared_from_this<Storage> {
  static constexpr size_t kPage = size_t(1) << 16;
  size_t bytes = 0;
  int64_t elements = 0;
  DType dtype = DType::F64;
  std::optional<Scalar> pending_fill;  // fill not yet issued on the device
  // A fused elementwise map not yet run, whose result is this buffer.
  // Operator results are left pending so that the next operator fuses them
  // into its own kernel: `(a + b.tile(n, 1)) * c` runs as one map.  The map
  // runs at the first use that does not fuse it.  Every storage it reads
  // lists it among its `readers`, and a write to one of them runs it first,
  // so the result is what the operator would have computed at once.  Only
  // programs that cannot raise (no checked Int32 arithmetic) are left
  // pending, so no error is raised later than the operator.
  std::shared_ptr<PendingMap> pending_map;
  std::vector<std::weak_ptr<Storage>> readers;
  void run_pending_map();
  // Runs the pending maps that read this buffer, before it is written.
  void before_write();
  ~Storage();
  void *dptr = nullptr;  // device memory: reach it through dev_read / dev_write
  // The device copy, current: the pending fill issued, dirty pages written back.
  const void *dev_read();
  // The same, for a use that may write: the host mirror is dropped.
  void *dev_write();
  // Host mirror bytes [off, off + len), fetched page by page as needed; the
  // write form marks the pages dirty.
  const unsigned char *host_read(size_t off, size_t len);
  unsigned char *host_write(size_t off, size_t len);
  // No host page newer than the device.
  bool host_clean() const { return dirty_ == 0; }
  // Drops the (clean) host mirror after a device write made outside
  // dev_write, which the storage's owner performs itself.
  void forget_host();

 private:
  std::vector<unsigned char> host_;
  std::vector<uint8_t> page_;  // per page: 0 not held, 1 held (equal to the device), 2 dirty
  size_t held_ = 0, dirty_ = 0;
  void ensure_pages(size_t off, size_t len);
};

// Host-mirror traffic since the process started (pages fetched from the
// device, runs of dirty pages written back, and their bytes).
struct MirrorStats {
  int64_t pages_fetched = 0, bytes_fetched = 0, runs_written = 0, bytes_written = 0;
};
MirrorStats mirror_stats();

// Coordinate transforms of a view chain (src/view_util/transforms.cr).
struct Transform {
  enum Kind { Region, Reshape, Permute, Reverse, Tile } kind;
  IndexRegion region;          // Region
  std::vector<int64_t> a, b;   // Reshape: src_shape, new_shape; Permute: pattern;
                               // Reverse: shape; Tile: src_shape, counts
};

class View;
class MutableView;
class ProcView;

// A view whose transform chain is longer than one descriptor holds: the
// storage it reads, the descriptor of the chain's oldest transforms (a view
// of shape mid_shape) and the newer ones (newest first).  Enough to rebuild
// the buffer index of each of its elements (ph_view_index, then gathers
// through the newer transforms).
struct DeepChain {
  std::shared_ptr<Storage> storage;
  ph_view_desc desc;
  Shape mid_shape;
  std::vector<Transform> tail;
};

// An operand of a fused elementwise kernel: storage + descriptor (+ the
// ProcView block applied to it, as an expression over input 0).  An operand
// of a view past one descriptor is a gathered copy of it (read when the view
// is used) and keeps the chain in `deep`: a write to the storage the chain
// reads then reads it element by element in the loop's order instead.
struct Operand {
  std::shared_ptr<Storage> storage;
  ph_view_desc desc;
  Shape shape;
  std::optional<Expr> body;
  std::shared_ptr<const DeepChain> deep;
  DType value_dtype() const;
};

// A pending map: its (not pending) inputs, its body over them, and the
// contiguous descriptor of its result.
struct PendingMap {
  std::vector<Operand> inputs;
  Expr body;
  ph_view_desc out;
  int spliced = 0;  // consumers it has been spliced into so far
};

// Device-resident NArray (src/n_array.cr): row-major buffer + shape.
class NArray {
 public:
  std::shared_ptr<Storage> storage;
  Shape shape_;

  NArray() = default;
  // NArray.fill(shape, value) (:230-232)
  static NArray fill(const Shape &shape, const Scalar &value, std::optional<DType> dtype = {});
  // NArray.of_buffer(shape, buffer) (:76-79), from host memory
  static NArray of_buffer(const Shape &shape, DType dtype, const void *host, size_t bytes);
  // Seeded synthetic data (counter-based splitmix64; see ph_fill_random)
  static NArray random(const Shape &shape, DType dtype, uint64_t seed, uint64_t offset, int dist,
                       double a, double b);
  static NArray empty(const Shape &shape, DType dtype);  // uninitialised

  const Shape &shape() const { return shape_; }
  int dimensions() const { return (int)shape_.size(); }
  int64_t size() const { return ShapeUtil::shape_to_size(shape_); }
  DType dtype() const { return storage->dtype; }
  bool empty_() const { return size() == 0; }
  bool scalar_() const { return size() == 1; }
  void *device_ptr() const;  // materialises a pending fill
  ph_view_desc desc() const;
  Operand operand() const;

  void to_host(void *out, size_t bytes) const;
  Scalar get(const Coord &coord) const;           // MultiIndexable#get
  void set_element(const Coord &coord, const Scalar &v);  // MultiWritable#set_element

  // MultiIndexable
  NArray get_chunk(const RegionLiteral &lit, bool drop = true) const;  // [] / get_chunk (:354-356)
  NArray get_chunk(const IndexRegion &region) const;                   // (:338-349)
  NArray unsafe_fetch_chunk(const IndexRegion &region) const;          // (n_array.cr:450-453)
  NArray get_available(const RegionLiteral &lit, bool drop = true) const;  // (:411-413)
  View view(const RegionLiteral &lit) const;                           // (:881-887)
  View view() const;
  MutableView mutable_view(const RegionLiteral &lit) const;            // (mutable.cr:6-12)
  MutableView mutable_view() const;
  ProcView process(const Expr &body) const;                            // (:889-895)
  View tile(const std::vector<int64_t> &counts) const;                 // lazy tile (:818-827)
  NArray reshape(const Shape &shape) const;                            // aliases (n_array.cr:429-433)
  NArray flatten() const;
  NArray permute(const std::vector<int64_t> *order = nullptr) const;   // (:795-803)
  NArray reverse() const;
  NArray clone() const;
  std::vector<NArray> slices(int axis = 0) const;                      // (:784-786)
  // NArray.concatenate(*narrs, axis) (n_array.cr:723-725, 727-751) and the
  // instance forms concatenate / concatenate! / push (:712-721, :688-710)
  static NArray concatenate(const std::vector<Operand> &narrs, int64_t axis = 0);
  NArray concatenate_with(const std::vector<Operand> &others, int64_t axis = 0) const;
  NArray &concatenate_inplace(const std::vector<Operand> &others, int64_t axis = 0);
  NArray &push(const std::vector<Operand> &others);

  // MultiWritable
  void set_chunk(const RegionLiteral &lit, const Operand &src);        // (multi_writable.cr:55-63)
  void set_chunk(const RegionLiteral &lit, const Scalar &value);       // (:67-70)
  void unsafe_set_chunk(const IndexRegion &region, const Operand &src);
  void unsafe_set_chunk(const IndexRegion &region, const Scalar &value);
  void set_mask(const Operand &mask, const Scalar &value);             // (n_array.cr:510-551)
  void set_mask(const Operand &mask, const Operand &src);

  // in-place map (map!, n_array.cr:622-628)
  NArray &map_inplace(const std::function<Expr(const Expr &)> &fn);
  // map_with_index / map_with_index! (n_array.cr:605-611, :632-638): the block
  // sees each element and its lexicographic index (Int32).
  NArray map_with_index(const std::function<Expr(const Expr &, const Expr &)> &fn) const;
  NArray &map_with_index_inplace(const std::function<Expr(const Expr &, const Expr &)> &fn);
  // map_with_coord! (n_array.cr:642-650): element, coordinate, index.
  NArray &map_with_coord_inplace(
      const std::function<Expr(const Expr &, const std::vector<Expr> &, const Expr &)> &fn);
};

// Lazy view (src/view.cr): source storage + a transform chain, folded into a
// ph_view_desc.
class View {
 public:
  std::shared_ptr<Storage> storage;
  Shape src_shape;
  Shape shape_;
  std::vector<Transform> transforms;  // newest first (compose! unshifts)
  // desc_ folds the oldest transforms, up to the descriptor's room for
  // reshape/tile stages; the newest tail_ transforms (ComposedTransform has
  // no bound, src/view_util/transforms.cr:28-101) are applied, when the view
  // is used, to a gathered copy of the view desc_ describes (of shape
  // mid_shape_).
  ph_view_desc desc_;
  size_t tail_ = 0;
  Shape mid_shape_;

  const Shape &shape() const { return shape_; }
  int dimensions() const { return (int)shape_.size(); }
  int64_t size() const { return ShapeUtil::shape_to_size(shape_); }
  DType dtype() const { return storage->dtype; }
  Operand operand() const;
  const ph_view_desc &desc() const { return desc_; }

  View view(const RegionLiteral &lit) const;
  View &restrict_to(const RegionLiteral &lit);   // (:43-49)
  View &restrict_to(const IndexRegion &region);  // (:51-56)
  View &reshape_(const Shape &shape);            // reshape! (:58-66)
  View &permute_(const std::vector<int64_t> *order);  // permute! (:72-81)
  View &reverse_();                              // reverse! (:96-99)
  View &tile_(const std::vector<int64_t> &counts);
  View reshape(const Shape &shape) const;
  View permute(const std::vector<int64_t> *order = nullptr) const;
  View reverse() const;
  NArray to_narr() const;                        // (:123-126)
  ProcView process(const Expr &body) const;      // (:119-121)
  Coord apply(const Coord &coord) const;         // ComposedTransform#apply
  Scalar get(const Coord &coord) const;
  bool deep() const { return tail_ > 0; }

 protected:
  void push(const Transform &t, const Shape &new_shape, const std::function<bool(ph_view_desc &)> &compose);
  void replay_tail(View &on) const;
  int64_t buffer_index(const Coord &coord) const;
};

class MutableView : public View {
 public:
  MutableView mutable_view(const RegionLiteral &lit) const;
  void set_chunk(const RegionLiteral &lit, const Operand &src);   // through the transform
  void set_chunk(const RegionLiteral &lit, const Scalar &value);
  void unsafe_set_chunk(const IndexRegion &region, const Operand &src);
  void set_element(const Coord &coord, const Scalar &v);

 private:
  void deep_write(const std::function<void(MutableView &)> &fn);
};

// ProcView (src/proc_view.cr): a view plus one element block.
class ProcView : public View {
 public:
  Expr body;  // over input 0 = the view's element
  Operand operand() const;
  NArray to_narr() const;
  ProcView process(const Expr &next) const;  // composition new(proc(x))
  Scalar get(const Coord &coord) const;
};

// --- elementwise (operator macros, src/multi_indexable.cr:931-985) ---------
// Any MultiIndexable operand: NArray, View, ProcView (the block is fused into
// the kernel) or a scalar.
NArray elementwise(const std::string &op, const Operand &a, const Operand &b);
NArray elementwise(const std::string &op, const Operand &a, const Scalar &b);
NArray elementwise(const std::string &op, const Scalar &a, const Operand &b);  // patches/number.cr
NArray elementwise_unary(const std::string &op, const Operand &a);
NArray map(const Operand &a, const std::function<Expr(const Expr &)> &fn);
// map_with_coord (src/multi_indexable.cr:708-712, NArray src/n_array.cr:614-
// 618): the block sees each element, its coordinate (Int32 per axis) and its
// lexicographic index.
NArray map_with_coord(const Operand &a,
                      const std::function<Expr(const Expr &, const std::vector<Expr> &, const Expr &)> &fn);
// The coordinate of the element at lexicographic index `idx` in `shape`, one
// Int32 expression per axis.
std::vector<Expr> coord_exprs(const Shape &shape, const Expr &idx);
// map_with over several operands of one shape (src/multi_indexable.cr:1026-1078)
NArray map_with(const std::vector<Operand> &ops, const std::function<Expr(const std::vector<Expr> &)> &fn);
// eq (:899-913): elementwise equality, DimensionError on shape mismatch
NArray eq(const Operand &a, const Operand &b);
// equals? with == (:871-879)
bool equals(const Operand &a, const Operand &b);

// --- reductions -------------------------------------------------------------
enum class Acc { F32C = PH_ACC_F32C, F64 = PH_ACC_F64 };
// Enumerable#sum / #max over `each` (lex order); Int32 sums raise OverflowError
Scalar sum(const Operand &a, Acc acc = Acc::F32C);
Scalar max(const Operand &a);
// slices(axis).map(&.sum) / (&.max)
std::vector<double> slices_sum(const Operand &a, int axis, Acc acc = Acc::F32C);
std::vector<double> slices_max(const Operand &a, int axis);
struct Reduce2D {
  double full_sum, full_max;
  std::vector<double> row_sum, row_max, col_sum, col_max;
};
// One pass over a 2-D operand: sum and max in full, along slices(0) and slices(1).
Reduce2D reduce2d_all(const Operand &a, Acc acc = Acc::F32C);
// The same, leaving the results in device memory: a Float64 NArray laid out
// [full_sum, full_max, row_sum[R], row_max[R], col_sum[C], col_max[C]].
NArray reduce2d_all_device(const Operand &a, Acc acc = Acc::F32C);

// Device scratch owned by the caller (e.g. results of reductions).
struct DeviceDoubles {
  double *ptr = nullptr;
  int64_t n = 0;
  explicit DeviceDoubles(int64_t n);
  ~DeviceDoubles();
  std::vector<double> to_host() const;
};

// Explicit 2-D heat equation on the device (examples/heat_equation.cr
// generalised to 2-D, Dirichlet zero): steps in place on `u`; returns the FP64
// energies at 0, every, 2*every, ... (empty if every == 0).
std::vector<double> heat2d(NArray &u, double r, int64_t steps, int64_t every);

// A batch of small Float32 arrays (rank 1..3) packed into one device buffer.
// run() computes, for every array k, what the host loop
//     y_k = x_k.map &.+(s_k); sums[k] = x_k.sum; z_k = x_k.view(0..2.., ...).to_narr
// computes, in one persistent launch (mode 0 replays a HIP graph; 1 launches
// directly; 2 is the unbatched three-launches-per-array form).
class Batch {
 public:
  Batch(const std::vector<Shape> &shapes, const std::vector<float> &scalars);
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  int64_t count() const { return (int64_t)shapes.size(); }
  // x as array k (a copy), y_k, z_k (copies)
  NArray x_at(int64_t k) const;
  NArray y_at(int64_t k) const;
  NArray z_at(int64_t k) const;
  // x (whole packed buffer, pads included) <- U[a, b) draws of stream `seed`
  void fill_uniform(uint64_t seed, double a, double b);
  void run(int mode = 0);
  // (work-item elements, workgroups, queue 1 / static 0)
  std::tuple<int64_t, int, int> tune();
  void schedule(int64_t chunk, int grid, int dynamic);
  std::vector<double> sums() const;

  std::vector<Shape> shapes;
  std::vector<float> scalars;
  std::vector<int64_t> offsets_x, offsets_z;
  NArray x, y, z;  // packed, flat Float32
  std::unique_ptr<DeviceDoubles> sums_dev;

 private:
  void prepare_run();
  void *handle_ = nullptr;
  NArray piece(const NArray &flat, int64_t off, const Shape &shape) const;
};

// One record of a general batch (ph_gbatch_*): array k's shape, the indices
// of its map program (results into y) and reduction program (its sum and /
// or maximum), the region copied into z (range literals, as `view` takes
// them, canonicalised against the shape), and its scalar (program input 1).
struct GBatchRecord {
  Shape shape;
  int map = -1;
  int red = -1;
  int red_ops = 0;  // PH_RED_SUM | PH_RED_MAX
  bool slice = false;
  RegionLiteral region;
  double scalar = 0.0;
};

// A general batch of small arrays packed into one device buffer: for each
// record the work `x_k.map { |v| map(v, s_k) }`, `x_k.map { |v| red(v, s_k)
// }.sum / .max` and `x_k.view(region).to_narr` a host loop would do, in one
// persistent launch (mode 0 replays a HIP graph; 1 launches directly; 2
// launches the kernel once per array; 3 runs each record through the
// unbatched entry points).  Programs are element
// expressions over input 0 (the element, x's dtype) and input 1 (the
// record's scalar, x's dtype).
class GBatch {
 public:
  GBatch(DType x_dtype, DType y_dtype, const std::vector<Expr> &progs, const std::vector<GBatchRecord> &records);
  ~GBatch();
  GBatch(const GBatch &) = delete;
  GBatch &operator=(const GBatch &) = delete;

  int64_t count() const { return (int64_t)records.size(); }
  NArray x_at(int64_t k) const;  // copies of x_k, y_k, z_k
  NArray y_at(int64_t k) const;
  NArray z_at(int64_t k) const;
  // copies a flat array of x's dtype and size (pads included) into x
  void set_x(const NArray &flat);
  void run(int mode = 0);
  // (work-item elements, workgroups, queue 1 / static 0)
  std::tuple<int64_t, int, int> tune();
  std::vector<double> sums() const;
  std::vector<double> maxes() const;

  DType x_dtype, y_dtype;
  std::vector<GBatchRecord> records;
  std::vector<Shape> z_shapes;  // the regions' shapes (degenerate axes dropped)
  std::vector<int64_t> offsets_xy, offsets_z;
  NArray x, y, z;  // packed, flat
  std::unique_ptr<DeviceDoubles> sums_dev, maxes_dev;

 private:
  void *handle_ = nullptr;
  void prepare_run();
  NArray piece(const NArray &flat, DType dt, int64_t off, const Shape &shape) const;
};

// Serialisation (n_array.cr:807-911): {"shape": [...], "elements": [...]}
// as JSON and as YAML, elements in buffer order, each as its type's to_s.
std::string to_json(const NArray &a);
NArray narray_from_json(const std::string &text, DType dtype);
std::string to_yaml(const NArray &a);
NArray narray_from_yaml(const std::string &text, DType dtype);
// Float#to_s: shortest round-trip digits, plain or d.ddde+XX notation.
std::string float_to_s(double v, bool single);

// Initialises the device (idempotent).
void init(int device = 0);
bool device_available();

}  // namespace phase
