# Phase::HIP — device-resident NArrays for ph-core on one MI355X, over
# LibPhHip (lib_ph_hip.cr).  Validation stays where ph-core does it
# (RangeSyntax, IndexRegion, ShapeUtil), so exceptions and messages are
# ph-core's own; the library is called with canonical regions only.
#
# Not compiled in this repository (no Crystal compiler in its image); the
# same design, compiled and tested, is phase::NArray in ../host/phase.cpp.
require "./lib_ph_hip"

module Phase::HIP
  @@ready = false

  # ph_init on `device` (src/n_array.cr:20 owns a host Slice; this owns a
  # device buffer instead).
  def self.init(device = 0) : Nil
    return if @@ready
    check(LibPhHip.init(device), "ph_init")
    @@ready = true
  end

  def self.check(code : Int32, what : String) : Nil
    return if code == 0
    msg = "#{what}: #{String.new(LibPhHip.last_error)}"
    case code
    when -1 then raise IndexError.new(msg)
    when -2 then raise ShapeError.new(msg)
    when -3 then raise DimensionError.new(msg)
    when -4 then raise ArgumentError.new(msg)
    when -5 then raise OverflowError.new(msg)
    when -6 then raise DivisionByZeroError.new(msg)
    else         raise Exception.new(msg)
    end
  end

  # Flags of ph_map / ph_reduce: what Crystal's checked arithmetic raises.
  def self.check_flags(flags : Int32, what : String) : Nil
    raise ArgumentError.new("Cannot raise an integer to a negative integer power, use floats for that") if flags & LibPhHip::FLAG_NEGEXP != 0
    raise DivisionByZeroError.new("Division by 0 (#{what})") if flags & LibPhHip::FLAG_DIVZERO != 0
    raise OverflowError.new("Arithmetic overflow (#{what})") if flags & LibPhHip::FLAG_OVERFLOW != 0
  end

  def self.dtype(t : Int32.class) : LibPhHip::DType
    LibPhHip::DType::I32
  end

  def self.dtype(t : Float32.class) : LibPhHip::DType
    LibPhHip::DType::F32
  end

  def self.dtype(t : Float64.class) : LibPhHip::DType
    LibPhHip::DType::F64
  end

  def self.dtype(t : Bool.class) : LibPhHip::DType
    LibPhHip::DType::Bool
  end

  # Row-major descriptor of a whole buffer of `shape`.
  def self.contiguous(dt : LibPhHip::DType, shape : Array(Int32)) : LibPhHip::ViewDesc
    d = LibPhHip::ViewDesc.new
    d.dtype = dt.value
    d.extent = shape.reduce(1_i64) { |acc, e| acc * e }
    v = d.view
    v.rank = shape.size
    shp = v.shape
    str = v.stride
    st = 1_i64
    (shape.size - 1).downto(0) do |i|
      shp[i] = shape[i].to_i64
      str[i] = st
      st *= shape[i]
    end
    v.shape = shp
    v.stride = str
    d.view = v
    d
  end

  # The descriptor of `region` (canonical, IndexRegion#local_to_absolute_unsafe,
  # src/index_region.cr:621-637) over a row-major buffer of `shape`: each
  # region axis is an affine axis of the buffer, dropped axes fold into the
  # offset (IndexRegion.compute_reduced_shape, :323-335).
  def self.region_desc(dt : LibPhHip::DType, shape : Array(Int32), region : IndexRegion) : LibPhHip::ViewDesc
    d = contiguous(dt, shape)
    base = d.view
    first = region.@first
    step = region.@step
    proper = region.@proper_shape
    reduced = region.@reduced_shape
    v = LibPhHip::Affine.new
    shp = v.shape
    str = v.stride
    offset = 0_i64
    k = 0
    shape.size.times do |i|
      offset += first[i].to_i64 * base.stride[i]
      keep = !(region.drop && region.degeneracy[i]) || reduced.size == shape.size
      if keep
        shp[k] = proper[i].to_i64
        str[k] = step[i].to_i64 * base.stride[i]
        k += 1
      end
    end
    v.rank = k
    v.offset = offset
    v.shape = shp
    v.stride = str
    d.view = v
    d
  end

  # ---------------------------------------------------------------------------
  # Element programs.  A block is run once on Tracer values: each operator
  # records a node, and the node's Crystal type is computed by Crystal itself
  # (typeof of the same operation on zeros), so the program has exactly the
  # stdlib's promotions (Float32 op Float64 -> Float64, Int32 / Int32 ->
  # Float64, ...).  The tree is then lowered to at most 48 instructions over
  # 8 registers (the subtree needing more registers first).
  # ---------------------------------------------------------------------------
  enum Op : UInt8
    Load; Const; Cast; Add; Sub; Mul; WAdd; WSub; WMul; Div; FloorDiv; Mod; PowI; IPow; WIPow; Pow
    Neg; BNot; And; Or; Xor; Gt; Lt; Ge; Le; Eq; Ne; Cmp; Sqrt; Abs; Select; Index
  end

  class Node
    getter op : Op
    getter dt : LibPhHip::DType   # result type
    getter opdt : LibPhHip::DType # the type the operation computes in
    getter args : Array(Node)
    getter input : Int32
    getter imm_i : Int64
    getter imm_f : Float64

    def initialize(@op, @dt, @opdt, @args = [] of Node, @input = 0, @imm_i = 0_i64, @imm_f = 0.0)
    end

    # registers needed to evaluate this subtree
    def need : Int32
      return 1 if @args.empty?
      ns = @args.map(&.need)
      return ns[0] if ns.size == 1
      m = ns.max
      ns[0] == ns[1] ? m + 1 : m
    end
  end

  def self.cast(n : Node, to : LibPhHip::DType) : Node
    return n if n.dt == to
    Node.new(Op::Cast, to, n.dt, [n])
  end

  struct Tracer(T)
    getter node : Node

    def initialize(@node)
    end

    def self.input(k : Int32, storage : LibPhHip::DType) : self
      # BF16 storage widens to Float32 on load
      dt = storage == LibPhHip::DType::BF16 ? LibPhHip::DType::F32 : storage
      new(Node.new(Op::Load, dt, dt, input: k))
    end

    def self.const(v : T) : self
      dt = HIP.dtype(T)
      if v.is_a?(Float)
        new(Node.new(Op::Const, dt, dt, imm_f: v.to_f64))
      else
        new(Node.new(Op::Const, dt, dt, imm_i: (v == true ? 1 : v == false ? 0 : v.to_i64)))
      end
    end

    private def self.wrap(v : Tracer) : Tracer
      v
    end

    private def self.wrap(v : Number | Bool)
      Tracer(typeof(v)).const(v)
    end

    # arithmetic: the result type is what Crystal gives for the same operands
    {% for name, op in {"+" => "Add", "-" => "Sub", "*" => "Mul"} %}
      def {{name.id}}(other)
        o = Tracer.wrap(other)
        r = HIP.dtype(typeof(T.zero {{name.id}} o.zero_value))
        Tracer(typeof(T.zero {{name.id}} o.zero_value)).new(
          Node.new(Op::{{op.id}}, r, r, [HIP.cast(node, r), HIP.cast(o.node, r)]))
      end
    {% end %}

    {% for name, op in {"&+" => "WAdd", "&-" => "WSub", "&*" => "WMul"} %}
      def {{name.id}}(other)
        o = Tracer.wrap(other)
        Tracer(Int32).new(Node.new(Op::{{op.id}}, LibPhHip::DType::I32, LibPhHip::DType::I32, [node, o.node]))
      end
    {% end %}

    def /(other)
      o = Tracer.wrap(other)
      r = HIP.dtype(typeof(T.zero / o.zero_value))
      Tracer(typeof(T.zero / o.zero_value)).new(Node.new(Op::Div, r, r, [HIP.cast(node, r), HIP.cast(o.node, r)]))
    end

    {% for name, op in {"//" => "FloorDiv", "%" => "Mod"} %}
      def {{name.id}}(other)
        o = Tracer.wrap(other)
        r = HIP.dtype(typeof(T.zero {{name.id}} o.zero_value))
        Tracer(typeof(T.zero {{name.id}} o.zero_value)).new(
          Node.new(Op::{{op.id}}, r, r, [HIP.cast(node, r), HIP.cast(o.node, r)]))
      end
    {% end %}

    def **(other)
      o = Tracer.wrap(other)
      r = HIP.dtype(typeof(T.zero ** o.zero_value))
      op = if r == LibPhHip::DType::I32
             Op::IPow
           elsif o.node.dt == LibPhHip::DType::I32
             Op::PowI
           else
             Op::Pow
           end
      b = op == Op::Pow ? HIP.cast(o.node, r) : o.node
      Tracer(typeof(T.zero ** o.zero_value)).new(Node.new(op, r, r, [HIP.cast(node, r), b]))
    end

    {% for name, op in {">" => "Gt", "<" => "Lt", ">=" => "Ge", "<=" => "Le", "==" => "Eq", "!=" => "Ne"} %}
      def {{name.id}}(other)
        o = Tracer.wrap(other)
        c = HIP.dtype(typeof(T.zero + o.zero_value))
        Tracer(Bool).new(Node.new(Op::{{op.id}}, LibPhHip::DType::Bool, c, [HIP.cast(node, c), HIP.cast(o.node, c)]))
      end
    {% end %}

    def <=>(other)
      o = Tracer.wrap(other)
      c = HIP.dtype(typeof(T.zero + o.zero_value))
      Tracer(Int32).new(Node.new(Op::Cmp, LibPhHip::DType::I32, c, [HIP.cast(node, c), HIP.cast(o.node, c)]))
    end

    def - : Tracer(T)
      Tracer(T).new(Node.new(Op::Neg, node.dt, node.dt, [node]))
    end

    def abs : Tracer(T)
      Tracer(T).new(Node.new(Op::Abs, node.dt, node.dt, [node]))
    end

    def zero_value : T
      T.zero
    end
  end

  # Lowers a node tree to a LibPhHip::Prog over `n_in` inputs.
  class Lowering
    @prog = LibPhHip::Prog.new
    @used = StaticArray(Bool, 8).new(false)

    def self.compile(root : Node, n_in : Int32) : LibPhHip::Prog
      l = new
      r = l.lower(root)
      l.finish(r, root.dt, n_in)
    end

    def finish(r : Int32, dt : LibPhHip::DType, n_in : Int32) : LibPhHip::Prog
      @prog.out_reg = r
      @prog.out_dtype = dt.value
      @prog.n_in = n_in
      @prog
    end

    private def alloc : Int32
      8.times do |r|
        unless @used[r]
          @used[r] = true
          return r
        end
      end
      raise ArgumentError.new("element expression needs more than 8 registers")
    end

    private def emit(op : Op, dt : LibPhHip::DType, a = 0, b = 0, c = 0, imm_i = 0_i64, imm_f = 0.0) : Int32
      raise ArgumentError.new("element expression too long") if @prog.n_insn >= LibPhHip::MAX_INSN
      insn = LibPhHip::Insn.new
      insn.op = op.value
      insn.dt = dt.value.to_u8
      insn.a = a.to_u8
      insn.b = b.to_u8
      insn.c = c.to_u8
      imm = LibPhHip::Imm.new
      if op == Op::Const && (dt == LibPhHip::DType::F32 || dt == LibPhHip::DType::F64)
        imm.f = imm_f
      else
        imm.i = imm_i
      end
      insn.imm = imm
      dst = alloc
      insn.dst = dst.to_u8
      ins = @prog.insn
      ins[@prog.n_insn] = insn
      @prog.insn = ins
      @prog.n_insn += 1
      dst
    end

    def lower(n : Node) : Int32
      case n.op
      when Op::Load  then return emit(Op::Load, n.dt, a: n.input)
      when Op::Index then return emit(Op::Index, LibPhHip::DType::I32)
      when Op::Const then return emit(Op::Const, n.dt, imm_i: n.imm_i, imm_f: n.imm_f)
      end
      regs = [] of Int32
      if n.args.size >= 2 && n.args[1].need > n.args[0].need
        rb = lower(n.args[1])
        ra = lower(n.args[0])
        regs = [ra, rb]
        regs << lower(n.args[2]) if n.args.size == 3
      else
        n.args.each { |a| regs << lower(a) }
      end
      regs.each { |r| @used[r] = false }
      dt = n.op == Op::Cast || n.op == Op::PowI ? n.dt : n.opdt
      c = n.op == Op::Cast ? n.opdt.value : (regs[2]? || 0)
      emit(n.op, dt, regs[0], regs[1]? || 0, c)
    end
  end

  # ---------------------------------------------------------------------------
  # DArray(T): an NArray whose buffer lives on the device.  It includes
  # ph-core's own MultiIndexable::Mutable(T), so every inherited method keeps
  # working (through unsafe_fetch_element / unsafe_set_element, one element
  # at a time); the hot-path entry points are overridden with device calls.
  # ---------------------------------------------------------------------------
  class DArray(T)
    include MultiIndexable::Mutable(T)

    getter ptr : Void*
    @shape : Array(Int32)

    protected def initialize(@shape : Array(Int32))
      HIP.init
      p = Pointer(Void).null
      HIP.check(LibPhHip.alloc(pointerof(p), LibC::SizeT.new(size.to_i64 * sizeof(T))), "ph_alloc")
      @ptr = p
    end

    def finalize
      LibPhHip.free(@ptr)
    end

    def shape_internal : Array(Int32) # src/multi_indexable.cr:34
      @shape
    end

    def size : Int32
      @shape.reduce(1) { |acc, e| acc * e }
    end

    def desc : LibPhHip::ViewDesc
      HIP.contiguous(HIP.dtype(T), @shape)
    end

    # NArray.fill (src/n_array.cr:230-232)
    def self.fill(shape : Array(Int32), value : T) : self
      a = new(shape.clone)
      d = a.desc
      v = value
      HIP.check(LibPhHip.fill(a.ptr, pointerof(d), pointerof(v).as(Void*)), "NArray.fill") if a.size > 0
      a
    end

    # NArray.of_buffer (src/n_array.cr:76-79) from a host Slice
    def self.of_buffer(shape : Array(Int32), buffer : Slice(T)) : self
      a = new(shape.clone)
      HIP.check(LibPhHip.h2d(a.ptr, buffer.to_unsafe.as(Void*), LibC::SizeT.new(buffer.bytesize)), "ph_h2d")
      a
    end

    def self.from_narr(n : NArray(T)) : self
      of_buffer(n.shape, n.buffer)
    end

    def to_narr : NArray(T)
      buf = Slice(T).new(size)
      HIP.check(LibPhHip.d2h(buf.to_unsafe.as(Void*), @ptr, LibC::SizeT.new(buf.bytesize)), "ph_d2h")
      NArray(T).of_buffer(@shape.clone, buf)
    end

    # element access: one element copied each way (phase::Storage in C++
    # keeps a paged host mirror instead)
    def unsafe_fetch_element(coord) : T
      v = uninitialized T
      HIP.check(LibPhHip.d2h(pointerof(v).as(Void*), (@ptr.as(T*) + index_of(coord)).as(Void*),
        LibC::SizeT.new(sizeof(T))), "ph_d2h")
      v
    end

    def unsafe_set_element(coord, value : T)
      v = value
      HIP.check(LibPhHip.h2d((@ptr.as(T*) + index_of(coord)).as(Void*), pointerof(v).as(Void*),
        LibC::SizeT.new(sizeof(T))), "ph_h2d")
    end

    private def index_of(coord) : Int64
      idx = 0_i64
      coord.each_with_index { |c, i| idx = idx * @shape[i] + c }
      idx
    end

    # NArray#unsafe_fetch_chunk (src/n_array.cr:450-453): one strided copy
    def unsafe_fetch_chunk(region : IndexRegion) : DArray(T)
      out = DArray(T).new(region.shape)
      return out if out.size == 0
      sd = HIP.region_desc(HIP.dtype(T), @shape, region)
      dd = out.desc
      HIP.check(LibPhHip.strided_copy(out.ptr, pointerof(dd), @ptr, pointerof(sd), 0), "gather")
      out
    end

    # NArray#unsafe_set_chunk (src/n_array.cr:484-500).  A source that is this
    # same array (another region of it) needs nothing special: the library
    # follows ph-core's element order when source and destination share a
    # buffer.
    def unsafe_set_chunk(region : IndexRegion, src : DArray(T))
      dd = HIP.region_desc(HIP.dtype(T), @shape, region)
      sd = HIP.contiguous(HIP.dtype(T), dd.view.shape.to_a[0, dd.view.rank].map(&.to_i32))
      sd.extent = src.size.to_i64
      HIP.check(LibPhHip.strided_copy(@ptr, pointerof(dd), src.ptr, pointerof(sd), 0), "scatter")
    end

    def unsafe_set_chunk(region : IndexRegion, value : T)
      dd = HIP.region_desc(HIP.dtype(T), @shape, region)
      v = value
      HIP.check(LibPhHip.fill(@ptr, pointerof(dd), pointerof(v).as(Void*)), "set_chunk")
    end

    # One fused map over this array and `others` (same shape): the block is
    # recorded once on Tracer values and compiled by the library.
    # (MultiIndexable#map_with, src/multi_indexable.cr:1026-1102; NArray#map,
    # src/n_array.cr:589-595)
    def map_with(*others : DArray, &)
      ins = [@ptr] + others.to_a.map(&.ptr)
      descs = [desc] + others.to_a.map(&.desc)
      others.each do |o|
        raise ShapeError.new("Could not simultaneously map MultiIndexables with shapes #{o.shape} and #{@shape}.") if o.shape_internal != @shape
      end
      t = yield Tracer(T).input(0, HIP.dtype(T)), *others.map_with_index { |o, k| typeof(o).trace_input(k + 1) }
      run_map(t, ins, descs)
    end

    def map(&)
      t = yield Tracer(T).input(0, HIP.dtype(T))
      run_map(t, [@ptr], [desc])
    end

    protected def self.trace_input(k : Int32)
      Tracer(T).input(k, HIP.dtype(T))
    end

    private def run_map(t : Tracer(U), ins : Array(Void*), descs : Array(LibPhHip::ViewDesc)) forall U
      prog = Lowering.compile(t.node, ins.size)
      out = DArray(U).new(@shape.clone)
      return out if out.size == 0
      od = out.desc
      flags = 0
      HIP.check(LibPhHip.map(pointerof(prog), ins.size, ins.to_unsafe, descs.to_unsafe, out.ptr, pointerof(od),
        pointerof(flags)), "map")
      HIP.check_flags(flags, "elementwise")
      out
    end

    # The operator macros (src/multi_indexable.cr:931-985): equal shapes,
    # else ShapeError; a scalar operand is a constant of the program.
    {% for name in %w(+ - * / // % ** &+ &- &* > < >= <= <=>) %}
      def {{name.id}}(other : DArray)
        if other.shape_internal != @shape
          raise ShapeError.new("The shape of this MultiIndexable (#{@shape}) does not match the shape of the one provided (#{other.shape_internal}), so '{{name.id}}' cannot be applied element-wise.")
        end
        map_with(other) { |a, b| a {{name.id}} b }
      end

      def {{name.id}}(other : Number)
        map { |a| a {{name.id}} other }
      end
    {% end %}

    # NArray#[]=(mask, value) (src/n_array.cr:510-536): a select program
    # writing into this array; a mask or source that reads this array through
    # another layout sees the loop's earlier writes, as in ph-core.
    def []=(mask : DArray(Bool), value : T)
      raise DimensionError.new("Cannot perform masking: mask shape does not match array shape.") if mask.shape_internal != @shape
      m = Tracer(Bool).input(0, LibPhHip::DType::Bool)
      own = Tracer(T).input(1, HIP.dtype(T))
      c = Tracer(T).const(value)
      sel = Node.new(Op::Select, own.node.dt, own.node.dt, [c.node, own.node, m.node])
      prog = Lowering.compile(sel, 2)
      ins = [mask.ptr, @ptr]
      descs = [mask.desc, desc]
      od = desc
      flags = 0
      HIP.check(LibPhHip.map(pointerof(prog), 2, ins.to_unsafe, descs.to_unsafe, @ptr, pointerof(od),
        pointerof(flags)), "mask")
    end

    # Enumerable#sum / #max (src/multi_indexable.cr:18) and
    # slices(k).map(&.sum) (:742-786): one fused reduction pass
    def sum : Float64
      reduce_all(LibPhHip::RED_SUM)[0]
    end

    def max : Float64
      raise EmptyError.new if size == 0
      reduce_all(LibPhHip::RED_MAX)[1]
    end

    def slices_sum(axis : Int32) : Array(Float64)
      out = DArray(Float64).new([@shape[axis]])
      prog = Lowering.compile(Tracer(T).input(0, HIP.dtype(T)).node, 1)
      d = desc
      ins = [@ptr]
      flags = 0
      HIP.check(LibPhHip.reduce(pointerof(prog), 1, ins.to_unsafe, pointerof(d), axis, LibPhHip::RED_SUM,
        LibPhHip::ACC_F32C, out.ptr.as(Float64*), Pointer(Float64).null, pointerof(flags)), "slices sum")
      HIP.check_flags(flags, "sum")
      out.to_narr.buffer.to_a
    end

    private def reduce_all(ops : Int32) : StaticArray(Float64, 2)
      res = DArray(Float64).new([2])
      prog = Lowering.compile(Tracer(T).input(0, HIP.dtype(T)).node, 1)
      d = desc
      ins = [@ptr]
      flags = 0
      HIP.check(LibPhHip.reduce(pointerof(prog), 1, ins.to_unsafe, pointerof(d), -1, ops, LibPhHip::ACC_F32C,
        res.ptr.as(Float64*), res.ptr.as(Float64*) + 1, pointerof(flags)), "reduce")
      HIP.check_flags(flags, "sum")
      b = res.to_narr.buffer
      StaticArray[b[0], b[1]]
    end
  end

  # examples/heat_equation.cr:38-51 generalised to 2-D: `steps` explicit
  # steps of the 5-point stencil on a DArray(Float64) or DArray(Float32)
  # grid, in place; returns the FP64 energy every `every` steps.
  def self.heat2d(u : DArray(F), r : Float64, steps : Int32, every = 0) : Array(Float64) forall F
    raise DimensionError.new("heat2d needs a 2-D grid") unless u.shape_internal.size == 2
    ny, nx = u.shape_internal
    tmp = DArray(F).new([ny, nx])
    n = every > 0 ? steps // every + 1 : 0
    e = DArray(Float64).new([n > 0 ? n : 1])
    check(LibPhHip.heat2d(dtype(F).value, u.ptr, tmp.ptr, ny.to_i64, nx.to_i64, r, steps.to_i64, every.to_i64,
      every > 0 ? e.ptr.as(Float64*) : Pointer(Float64).null), "heat2d")
    n > 0 ? e.to_narr.buffer.to_a[0, n] : [] of Float64
  end
end
