# Crystal binding of libphhip.so, the MI355X backend of ph-core's NArray path.
# Every entry point is declared in include/phhip.h, whose comments cite the
# ph-core code each one replaces.  This file and phase_hip.cr are the Crystal
# side of the drop-in; the image that built this repository has no Crystal
# compiler, so they are not compiled here (phase::* in ../host is the same
# integration in C++, and it is what the tests drive).
@[Link(ldflags: "-L#{__DIR__}/../lib -lphhip -Wl,-rpath,#{__DIR__}/../lib")]
lib LibPhHip
  MAX_RANK   =  8
  MAX_STAGES =  2
  MAX_INSN   = 48
  MAX_INPUTS =  8
  NREG       =  8

  enum DType : Int32
    I32  = 0
    F32  = 1
    F64  = 2
    BF16 = 3
    Bool = 4
  end

  # status codes: 0 ok; -1 IndexError, -2 ShapeError, -3 DimensionError,
  # -4 ArgumentError, -5 OverflowError, -6 DivisionByZeroError,
  # -10 HIP, -11 not initialised, -12 unsupported, -13 out of memory
  FLAG_OVERFLOW = 1
  FLAG_DIVZERO  = 2
  FLAG_NEGEXP   = 4

  RED_SUM  = 1
  RED_MAX  = 2
  ACC_F32C = 0
  ACC_F64  = 1

  struct Affine
    rank : Int32
    pad_ : Int32
    offset : Int64
    shape : StaticArray(Int64, 8)
    stride : StaticArray(Int64, 8)
  end

  struct ViewDesc
    dtype : Int32
    nstage : Int32
    extent : Int64
    view : Affine
    stage : StaticArray(Affine, 2)
  end

  union Imm
    i : Int64
    f : Float64
  end

  struct Insn
    op : UInt8
    dt : UInt8
    dst : UInt8
    a : UInt8
    b : UInt8
    c : UInt8
    pad_ : UInt16
    imm : Imm
  end

  struct Prog
    n_insn : Int32
    n_in : Int32
    out_reg : Int32
    out_dtype : Int32
    insn : StaticArray(Insn, 48)
  end

  fun init = ph_init(device : Int32) : Int32
  fun device_count = ph_device_count(n : Int32*) : Int32
  fun last_error = ph_last_error : UInt8*
  fun alloc = ph_alloc(ptr : Void**, bytes : LibC::SizeT) : Int32
  fun free = ph_free(ptr : Void*) : Int32
  fun h2d = ph_h2d(dst : Void*, src : Void*, bytes : LibC::SizeT) : Int32
  fun d2h = ph_d2h(dst : Void*, src : Void*, bytes : LibC::SizeT) : Int32
  fun d2d = ph_d2d(dst : Void*, src : Void*, bytes : LibC::SizeT) : Int32
  fun sync = ph_sync : Int32
  fun fill = ph_fill(buf : Void*, dst : ViewDesc*, value : Void*) : Int32
  fun strided_copy = ph_strided_copy(dst : Void*, dd : ViewDesc*, src : Void*, sd : ViewDesc*,
                                     flags : Int32) : Int32
  fun fill_scatter = ph_fill_scatter(dst : Void*, dtype : Int32, dst_shape : Int64*, rank : Int32,
                                     first : Int64*, step : Int64*, count : Int64*, axis : Int32*,
                                     src : Void*, sd : ViewDesc*, fill : Void*) : Int32
  fun map = ph_map(prog : Prog*, n_in : Int32, ins : Void**, in_descs : ViewDesc*, out : Void*,
                   out_desc : ViewDesc*, flags_out : Int32*) : Int32
  fun view_index = ph_view_index(out : Int64*, d : ViewDesc*) : Int32
  fun map_indexed = ph_map_indexed(prog : Prog*, n_in : Int32, ins : Void**, in_descs : ViewDesc*,
                                   in_index : Int64**, out : Void*, out_desc : ViewDesc*,
                                   out_index : Int64*, flags_out : Int32*) : Int32
  fun reduce = ph_reduce(prog : Prog*, n_in : Int32, ins : Void**, in_descs : ViewDesc*,
                         keep_axis : Int32, ops : Int32, acc : Int32, out_sum : Float64*,
                         out_max : Float64*, flags_out : Int32*) : Int32
  fun reduce2d_all = ph_reduce2d_all(prog : Prog*, n_in : Int32, ins : Void**, in_descs : ViewDesc*,
                                     acc : Int32, full_sum : Float64*, full_max : Float64*,
                                     row_sum : Float64*, row_max : Float64*, col_sum : Float64*,
                                     col_max : Float64*, flags_out : Int32*) : Int32
  fun heat2d = ph_heat2d(dtype : Int32, u : Void*, tmp : Void*, ny : Int64, nx : Int64, r : Float64,
                         steps : Int64, energy_every : Int64, energy_out : Float64*) : Int32
  # allocation-free variant (no ph-core counterpart): energy partials in a
  # fixed per-device array, for callers that cannot allocate device memory
  fun heat2d_noalloc = ph_heat2d_noalloc(dtype : Int32, u : Void*, tmp : Void*, ny : Int64, nx : Int64,
                                         r : Float64, steps : Int64, energy_every : Int64,
                                         energy_out : Float64*) : Int32
  fun l2_diff = ph_l2_diff(dtype_a : Int32, a : Void*, dtype_b : Int32, b : Void*, n : Int64,
                           out : Float64*) : Int32
  fun batch_create = ph_batch_create(handle : Void**, n : Int32, ranks : Int32*, shapes : Int64*,
                                     x : Float32*, scalars : Float32*, y : Float32*, z : Float32*,
                                     sums : Float64*) : Int32
  fun batch_sizes = ph_batch_sizes(handle : Void*, xy : Int64*, zn : Int64*, offsets_x : Int64*,
                                   offsets_z : Int64*) : Int32
  fun batch_bind = ph_batch_bind(handle : Void*, x : Float32*, y : Float32*, z : Float32*,
                                 sums : Float64*) : Int32
  fun batch_run = ph_batch_run(handle : Void*, mode : Int32) : Int32
  fun batch_tune = ph_batch_tune(handle : Void*, chunk : Int64*, grid : Int32*, dynamic : Int32*) : Int32
  fun batch_schedule = ph_batch_schedule(handle : Void*, chunk : Int64, grid : Int32, dynamic : Int32) : Int32
  fun batch_destroy = ph_batch_destroy(handle : Void*) : Int32

  # general batches: per record a map program, a reduction program and a region
  struct BatchRec
    rank : Int32
    map : Int32
    red : Int32
    red_ops : Int32
    slice : Int32
    pad_ : Int32
    shape : StaticArray(Int64, 3)
    first : StaticArray(Int64, 3)
    step : StaticArray(Int64, 3)
    count : StaticArray(Int64, 3)
    scalar : Float64
  end
  fun gbatch_create = ph_gbatch_create(handle : Void**, x_dtype : Int32, y_dtype : Int32, n_progs : Int32,
                                       progs : Prog*, n : Int32, recs : BatchRec*) : Int32
  fun gbatch_sizes = ph_gbatch_sizes(handle : Void*, x_elements : Int64*, z_elements : Int64*,
                                     offsets_xy : Int64*, offsets_z : Int64*) : Int32
  fun gbatch_run = ph_gbatch_run(handle : Void*, x : Void*, y : Void*, z : Void*, sums : Float64*,
                                 maxes : Float64*, mode : Int32, flags_out : Int32*) : Int32
  fun gbatch_tune = ph_gbatch_tune(handle : Void*, x : Void*, y : Void*, z : Void*, sums : Float64*,
                                   maxes : Float64*, chunk : Int64*, grid : Int32*, dynamic : Int32*) : Int32
  fun gbatch_destroy = ph_gbatch_destroy(handle : Void*) : Int32
  fun jit_compiles = ph_jit_compiles : LibC::Long
  fun launch_count = ph_launch_count : LibC::Long
end
