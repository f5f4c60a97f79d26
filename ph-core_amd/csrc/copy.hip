// Strided copies, fills and seeded random fills (hand-written, ahead-of-time).
//
// Layout of the work: a copy over a logical shape is first collapsed
// (ph::collapse) to `rows x cols` with the innermost axis as `cols`.  One
// 256-thread workgroup owns a tile of kTile consecutive columns of one row;
// lane t of the group handles columns tile0 + e*256 + t for e < EPT, so for
// every unrolled e the 64 lanes of a wavefront touch 64 consecutive elements
// of the destination (fully coalesced) and consecutive source elements
// `s_inner` apart (coalesced for |s_inner| <= 2 within 512-byte spans).  All
// EPT loads are issued before the stores, keeping EPT x 256 B per wavefront in
// flight.  Row offsets are decoded once per workgroup, in scalar registers.
#include <cstring>

#include "common.hpp"

namespace {

constexpr int kBlock = 256;

struct RowArgs {
  int64_t rows;
  int64_t cols;
  int64_t ntiles;
  int64_t nblocks;
  int32_t outer_rank;
  int32_t pad_;
  int64_t outer_shape[PH_MAX_RANK];
  int64_t d_off, s_off;
  int64_t d_outer[PH_MAX_RANK];
  int64_t s_outer[PH_MAX_RANK];
  int64_t d_inner, s_inner;
};

__device__ __forceinline__ void decode_row(const RowArgs &a, int64_t row, int64_t *d_base,
                                           int64_t *s_base) {
  int64_t d = a.d_off, s = a.s_off;
  for (int k = a.outer_rank - 1; k >= 0; --k) {
    int64_t n = a.outer_shape[k];
    int64_t q = row / n;
    int64_t c = row - q * n;
    row = q;
    d += c * a.d_outer[k];
    s += c * a.s_outer[k];
  }
  *d_base = d;
  *s_base = s;
}

// Index arithmetic runs in I (int32_t when every address of both operands fits
// in 31 bits, the common case: ph-core caps element counts at Int32::MAX), and
// a unit innermost stride is a compile-time constant (DCONT / SCONT), so the
// per-element work is one add and one address computation.
template <typename E, int EPT, typename I, bool DCONT, bool SCONT>
__global__ __launch_bounds__(kBlock) void k_copy_rows(E *__restrict__ dst,
                                                      const E *__restrict__ src, RowArgs a) {
  constexpr int64_t kTile = (int64_t)kBlock * EPT;
  const I cols = (I)a.cols;
  const I di = (I)a.d_inner, si = (I)a.s_inner;
  for (int64_t b = blockIdx.x; b < a.nblocks; b += gridDim.x) {
    const int64_t row = b / a.ntiles;
    const int64_t tile = b - row * a.ntiles;
    int64_t db64, sb64;
    decode_row(a, row, &db64, &sb64);
    const I db = (I)db64, sb = (I)sb64;
    const I c0 = (I)(tile * kTile) + (I)threadIdx.x;
    E v[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const I c = c0 + (I)(e * kBlock);
      if (c < cols) v[e] = src[sb + (SCONT ? c : c * si)];
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const I c = c0 + (I)(e * kBlock);
      if (c < cols) {
        if (DCONT) __builtin_nontemporal_store(v[e], dst + db + c);  // contiguous rows: streamed
        else dst[db + c * di] = v[e];
      }
    }
  }
}

template <typename E, int EPT, typename I, bool DCONT>
__global__ __launch_bounds__(kBlock) void k_fill_rows(E *__restrict__ dst, E value, RowArgs a) {
  constexpr int64_t kTile = (int64_t)kBlock * EPT;
  const I cols = (I)a.cols;
  const I di = (I)a.d_inner;
  for (int64_t b = blockIdx.x; b < a.nblocks; b += gridDim.x) {
    const int64_t row = b / a.ntiles;
    const int64_t tile = b - row * a.ntiles;
    int64_t db64, sb64;
    decode_row(a, row, &db64, &sb64);
    const I db = (I)db64;
    const I c0 = (I)(tile * kTile) + (I)threadIdx.x;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const I c = c0 + (I)(e * kBlock);
      if (c < cols) dst[db + (DCONT ? c : c * di)] = value;
    }
  }
}

// Both operands' addresses fit in 31 bits: [min, max] of every reachable
// index (the descriptors were validated against [0, extent)).
bool fits_i32(const ph_view_desc *a, const ph_view_desc *b) {
  return a->extent < (int64_t(1) << 31) && (!b || b->extent < (int64_t(1) << 31));
}

// Narrow rows (cols small): one element per thread, grid-stride, full decode.
template <typename E>
__global__ __launch_bounds__(kBlock) void k_copy_flat(E *__restrict__ dst,
                                                      const E *__restrict__ src, RowArgs a,
                                                      int64_t total, int has_src, E value) {
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * kBlock) {
    const int64_t row = g / a.cols;
    const int64_t c = g - row * a.cols;
    int64_t db, sb;
    decode_row(a, row, &db, &sb);
    dst[db + c * a.d_inner] = has_src ? src[sb + c * a.s_inner] : value;
  }
}

// General path: descriptors with reshape/tile stages.  Element g of the
// common logical shape is decoded into a coordinate and mapped through each
// descriptor in full.
__device__ __forceinline__ int64_t map_index(const ph_view_desc &d, const int64_t *coord) {
  int64_t idx = d.view.offset;
  for (int i = 0; i < d.view.rank; ++i) idx += coord[i] * d.view.stride[i];
  for (int s = 0; s < d.nstage; ++s) {
    const ph_affine &st = d.stage[s];
    int64_t out = st.offset;
    for (int j = st.rank - 1; j >= 0; --j) {
      const int64_t n = st.shape[j];
      const int64_t q = idx / n;
      out += (idx - q * n) * st.stride[j];
      idx = q;
    }
    idx = out;
  }
  return idx;
}

template <typename E>
__global__ __launch_bounds__(kBlock) void k_copy_general(E *__restrict__ dst, ph_view_desc dd,
                                                         const E *__restrict__ src,
                                                         ph_view_desc sd, int64_t total,
                                                         int has_src, E value) {
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * kBlock) {
    int64_t coord[PH_MAX_RANK];
    int64_t r = g;
    for (int i = dd.view.rank - 1; i >= 0; --i) {
      const int64_t n = dd.view.shape[i];
      const int64_t q = r / n;
      coord[i] = r - q * n;
      r = q;
    }
    const int64_t di = map_index(dd, coord);
    dst[di] = has_src ? src[map_index(sd, coord)] : value;
  }
}

// Buffer index of every coordinate of a view, lexicographic: out[g] = d(g).
__global__ __launch_bounds__(kBlock) void k_view_index(int64_t *__restrict__ out, ph_view_desc d, int64_t total) {
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total; g += (int64_t)gridDim.x * kBlock) {
    int64_t coord[PH_MAX_RANK];
    int64_t r = g;
    for (int i = d.view.rank - 1; i >= 0; --i) {
      const int64_t n = d.view.shape[i];
      const int64_t q = r / n;
      coord[i] = r - q * n;
      r = q;
    }
    out[g] = map_index(d, coord);
  }
}

// Transposing copy through LDS: the source is contiguous along axis A and the
// destination along another axis B (a permuted view materialised, or a
// contiguous array scattered into a permuted MutableView).  A workgroup owns
// a 64 x 64 tile of (A, B) for one coordinate of the other axes: it reads the
// tile with lanes along A (64 consecutive source elements per wavefront
// load), stages it in LDS (rows padded by one element, so the column reads
// below hit distinct banks), and writes it with lanes along B (64 consecutive
// destination elements per store).  Every element is read once and written
// once, each as part of whole cache lines.
struct TileArgs {
  int64_t na, nb;    // extents along A and B
  int64_t ta, tb;    // 64-element tiles along A and B
  int64_t nblocks;   // tiles in all
  int64_t s_a, s_b;  // source strides along A (+-1) and B
  int64_t d_a, d_b;  // destination strides along A and B (+-1)
  int64_t d_off, s_off;
  int32_t orank, pad_;
  int64_t o_shape[PH_MAX_RANK], o_dst[PH_MAX_RANK], o_src[PH_MAX_RANK];
};

constexpr int kTileT = 64;

template <typename E>
__global__ __launch_bounds__(kBlock) void k_copy_tiled(E *__restrict__ dst, const E *__restrict__ src,
                                                       TileArgs a) {
  constexpr int kRows = kTileT / (kBlock / 64);  // tile rows per wavefront
  __shared__ E lds[kTileT][kTileT + 1];          // [B][A]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t per = a.ta * a.tb;
  for (int64_t b = blockIdx.x; b < a.nblocks; b += gridDim.x) {
    int64_t o = b / per;
    const int64_t rem = b - o * per;
    const int64_t ib = rem / a.ta, ia = rem - ib * a.ta;
    int64_t db = a.d_off, sb = a.s_off;
    for (int k = a.orank - 1; k >= 0; --k) {
      const int64_t n = a.o_shape[k];
      const int64_t q = o / n;
      const int64_t c = o - q * n;
      o = q;
      db += c * a.o_dst[k];
      sb += c * a.o_src[k];
    }
    const int64_t a0 = ia * kTileT, b0 = ib * kTileT;
    // read: lane along A, wavefront w takes tile rows w, w + 4, ...
    const int64_t ar = a0 + lane;
    const E *sp = src + sb + ar * a.s_a + (b0 + wave) * a.s_b;
    E v[kRows];
#pragma unroll
    for (int j = 0; j < kRows; ++j) {
      v[j] = E{};
      if (ar < a.na && b0 + wave + 4 * j < a.nb) v[j] = sp[(int64_t)(4 * j) * a.s_b];
    }
#pragma unroll
    for (int j = 0; j < kRows; ++j) lds[wave + 4 * j][lane] = v[j];
    __syncthreads();
    // write: lane along B, wavefront w takes tile columns w, w + 4, ...
    const int64_t bw = b0 + lane;
    E *dp = dst + db + bw * a.d_b + (a0 + wave) * a.d_a;
#pragma unroll
    for (int j = 0; j < kRows; ++j)
      if (bw < a.nb && a0 + wave + 4 * j < a.na)
        __builtin_nontemporal_store(lds[lane][wave + 4 * j], dp + (int64_t)(4 * j) * a.d_a);
    __syncthreads();  // the next tile reuses the LDS
  }
}

// The same with 16-byte lanes on both sides, for 4- and 8-byte elements
// (VW = 4 or 2 per lane) and a T x T tile with T = 32 VW: a half-wavefront
// reads or writes a whole 512-byte run of a row, which suits DRAM pages
// better than the 256-byte runs above.  A lane loads a VW x VW block (VW
// consecutive source rows of VW consecutive A elements), transposes it in
// registers and stores VW vectors along B into LDS [A][B].  The 16-byte
// group index within an LDS row is XOR-swizzled by (a / VW) & 7, so eight
// consecutive lanes touch eight distinct bank groups on the stores and on
// the reads that follow.  Measured on 8192^2 Int32: 81 us back to back and
// 107 us with the source read cold, against 92 and 140 us for the 64 x 64
// tile.  Needs +1 strides along A in the source and B in the destination,
// both extents in whole vectors and 16-byte row starts.
template <typename E>
__global__ __launch_bounds__(kBlock) void k_copy_tiled_v(E *__restrict__ dst, const E *__restrict__ src,
                                                         TileArgs a) {
  typedef unsigned int V4 __attribute__((ext_vector_type(4)));
  constexpr int VW = 16 / (int)sizeof(E);  // elements per 16-byte lane: 4 or 2
  constexpr int T = 32 * VW;               // tile edge: 128 or 64
  static_assert(VW == 4 || VW == 2, "4- and 8-byte elements");
  __shared__ __align__(16) E lds[T * T];   // [A][B], 512-byte rows
  const int t = threadIdx.x;
  const int q = t & 31, r0 = t >> 5;       // lane in a row, row group 0..7
  const int64_t per = a.ta * a.tb;
  for (int64_t b = blockIdx.x; b < a.nblocks; b += gridDim.x) {
    int64_t o = b / per;
    const int64_t rem = b - o * per;
    const int64_t ib = rem / a.ta, ia = rem - ib * a.ta;
    int64_t db = a.d_off, sb = a.s_off;
    for (int k = a.orank - 1; k >= 0; --k) {
      const int64_t n = a.o_shape[k];
      const int64_t qq = o / n;
      const int64_t c = o - qq * n;
      o = qq;
      db += c * a.o_dst[k];
      sb += c * a.o_src[k];
    }
    const int64_t a0 = ia * T, b0 = ib * T;
    // read: row groups r = r0 + 8m (m < 4), rows VW r + jj, VW A elements per lane
    const int64_t ar = a0 + (int64_t)q * VW;
    V4 v[4][VW];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int jj = 0; jj < VW; ++jj) {
        v[m][jj] = V4{0u, 0u, 0u, 0u};
        const int64_t bv = b0 + VW * (r0 + 8 * m) + jj;
        if (ar < a.na && bv < a.nb) v[m][jj] = *(const V4 *)(src + sb + ar + bv * a.s_b);
      }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int g = ((r0 + 8 * m) ^ (q & 7)) * VW;  // swizzled group of row VW q + i
#pragma unroll
      for (int i = 0; i < VW; ++i) {
        V4 w;
        if constexpr (VW == 4) {
          w = V4{v[m][0][i], v[m][1][i], v[m][2][i], v[m][3][i]};
        } else {
          w = V4{v[m][0][2 * i], v[m][0][2 * i + 1], v[m][1][2 * i], v[m][1][2 * i + 1]};
        }
        *(V4 *)&lds[(VW * q + i) * T + g] = w;
      }
    }
    __syncthreads();
    // write: A rows r0 + 8j, VW B elements per lane
    const int64_t bw = b0 + (int64_t)q * VW;
#pragma unroll
    for (int j = 0; j < T / 8; ++j) {
      const int al = r0 + 8 * j;
      if (bw < a.nb && a0 + al < a.na) {
        const V4 w = *(const V4 *)&lds[al * T + ((q ^ ((al / VW) & 7)) * VW)];
        __builtin_nontemporal_store(w, (V4 *)(dst + db + bw + (a0 + al) * a.d_a));
      }
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}

// Whether k_copy_tiled_v applies: unit strides, extents in whole vectors and
// 16-byte row starts for every tile row of both operands.
bool tiled_vec_ok(const TileArgs &t, const void *dst, const void *src, int esize) {
  const int64_t vw = 16 / esize;
  if (t.s_a != 1 || t.d_b != 1 || t.na % vw != 0 || t.nb % vw != 0) return false;
  if ((uintptr_t)dst % 16 != 0 || (uintptr_t)src % 16 != 0) return false;
  auto al = [&](int64_t elems) { return (elems * esize) % 16 == 0; };
  if (!al(t.s_off) || !al(t.d_off) || !al(t.s_b) || !al(t.d_a)) return false;
  for (int k = 0; k < t.orank; ++k)
    if (!al(t.o_src[k]) || !al(t.o_dst[k])) return false;
  return true;
}

// The (A, B) axes of a transposing copy over the collapsed layout (operand 0
// the destination, 1 the source), or false when the innermost axis already
// serves both operands or either operand has no unit-stride axis.  Both
// extents must fill a good part of a tile.
bool tiled_axes(const ph::Layout &L, int *A, int *B) {
  if (L.rank < 2) return false;
  auto unit = [](int64_t s) { return s == 1 || s == -1; };
  const int q = L.rank - 1;
  if (unit(L.stride[0][q]) && unit(L.stride[1][q])) return false;
  *A = *B = -1;
  for (int k = 0; k < L.rank; ++k) {
    if (unit(L.stride[1][k]) && *A < 0) *A = k;
    if (unit(L.stride[0][k]) && *B < 0) *B = k;
  }
  return *A >= 0 && *B >= 0 && *A != *B && L.shape[*A] >= 16 && L.shape[*B] >= 16;
}

void make_tile_args(const ph::Layout &L, int A, int B, int tile, TileArgs *t) {
  std::memset(t, 0, sizeof(*t));
  t->na = L.shape[A];
  t->nb = L.shape[B];
  t->ta = (t->na + tile - 1) / tile;
  t->tb = (t->nb + tile - 1) / tile;
  t->s_a = L.stride[1][A];
  t->s_b = L.stride[1][B];
  t->d_a = L.stride[0][A];
  t->d_b = L.stride[0][B];
  t->d_off = L.offset[0];
  t->s_off = L.offset[1];
  int64_t others = 1;
  for (int k = 0; k < L.rank; ++k) {
    if (k == A || k == B) continue;
    t->o_shape[t->orank] = L.shape[k];
    t->o_dst[t->orank] = L.stride[0][k];
    t->o_src[t->orank] = L.stride[1][k];
    ++t->orank;
    others *= L.shape[k];
  }
  t->nblocks = others * t->ta * t->tb;
}

int64_t grid_for(int64_t blocks) {
  const int64_t cap = 1 << 20;
  return blocks < 1 ? 1 : (blocks > cap ? cap : blocks);
}

void make_row_args(const ph::Layout &L, int d_op, int s_op, int ept, RowArgs *a) {
  std::memset(a, 0, sizeof(*a));
  if (L.rank == 0) {
    a->rows = 1;
    a->cols = 1;
    a->d_off = L.offset[d_op];
    a->s_off = s_op >= 0 ? L.offset[s_op] : 0;
  } else {
    a->cols = L.shape[L.rank - 1];
    a->d_inner = L.stride[d_op][L.rank - 1];
    a->s_inner = s_op >= 0 ? L.stride[s_op][L.rank - 1] : 0;
    a->outer_rank = L.rank - 1;
    a->rows = 1;
    for (int k = 0; k < L.rank - 1; ++k) {
      a->outer_shape[k] = L.shape[k];
      a->rows *= L.shape[k];
      a->d_outer[k] = L.stride[d_op][k];
      a->s_outer[k] = s_op >= 0 ? L.stride[s_op][k] : 0;
    }
    a->d_off = L.offset[d_op];
    a->s_off = s_op >= 0 ? L.offset[s_op] : 0;
  }
  const int64_t tile = (int64_t)kBlock * ept;
  a->ntiles = (a->cols + tile - 1) / tile;
  a->nblocks = a->rows * a->ntiles;
}

template <typename E>
int launch_copy(void *dst, const ph_view_desc *dd, const void *src, const ph_view_desc *sd,
                const void *value_bits) {
  E value{};
  if (value_bits) std::memcpy(&value, value_bits, sizeof(E));
  const int64_t total = ph::desc_count(dd);
  if (total == 0) return PH_OK;
  int rc = ph::launch_gate("strided copy");
  if (rc) return rc;
  hipStream_t st = ph::stream();
  const bool staged = dd->nstage > 0 || (src && sd->nstage > 0);
  if (staged) {
    const ph_view_desc &s_desc = src ? *sd : *dd;
    hipLaunchKernelGGL(k_copy_general<E>, dim3((unsigned)grid_for((total + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, st, (E *)dst, *dd, (const E *)src, s_desc, total,
                       src ? 1 : 0, value);
    return ph::check_launch("k_copy_general");
  }
  const ph_view_desc *descs[2] = {dd, src ? sd : dd};
  ph::Layout L;
  ph::collapse(descs, 2, &L);
  constexpr int EPT = sizeof(E) >= 8 ? 4 : 8;
  int ax_a, ax_b;
  if (src && tiled_axes(L, &ax_a, &ax_b)) {
    TileArgs t;
    if constexpr (sizeof(E) >= 4) {
      make_tile_args(L, ax_a, ax_b, 32 * (16 / (int)sizeof(E)), &t);
      if (tiled_vec_ok(t, dst, src, (int)sizeof(E))) {
        hipLaunchKernelGGL((k_copy_tiled_v<E>), dim3((unsigned)grid_for(t.nblocks)), dim3(kBlock), 0, st,
                           (E *)dst, (const E *)src, t);
        return ph::check_launch("k_copy_tiled_v");
      }
    }
    make_tile_args(L, ax_a, ax_b, kTileT, &t);
    hipLaunchKernelGGL(k_copy_tiled<E>, dim3((unsigned)grid_for(t.nblocks)), dim3(kBlock), 0, st, (E *)dst,
                       (const E *)src, t);
    return ph::check_launch("k_copy_tiled");
  }
  RowArgs a;
  make_row_args(L, 0, src ? 1 : -1, EPT, &a);
  if (a.cols * 4 < (int64_t)kBlock * EPT && a.rows > 1) {
    hipLaunchKernelGGL(k_copy_flat<E>, dim3((unsigned)grid_for((total + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, st, (E *)dst, (const E *)src, a, total, src ? 1 : 0,
                       value);
    return ph::check_launch("k_copy_flat");
  }
  const dim3 grid((unsigned)grid_for(a.nblocks)), block(kBlock);
  const bool small = fits_i32(dd, src ? sd : nullptr);
  const bool dc = a.d_inner == 1, sc = a.s_inner == 1;
  if (src) {
#define PH_COPY_LAUNCH(I, D, S) \
  hipLaunchKernelGGL((k_copy_rows<E, EPT, I, D, S>), grid, block, 0, st, (E *)dst, (const E *)src, a)
    if (small) {
      if (dc && sc) PH_COPY_LAUNCH(int32_t, true, true);
      else if (dc) PH_COPY_LAUNCH(int32_t, true, false);
      else if (sc) PH_COPY_LAUNCH(int32_t, false, true);
      else PH_COPY_LAUNCH(int32_t, false, false);
    } else {
      PH_COPY_LAUNCH(int64_t, false, false);
    }
#undef PH_COPY_LAUNCH
    return ph::check_launch("k_copy_rows");
  }
  if (small && dc)
    hipLaunchKernelGGL((k_fill_rows<E, EPT, int32_t, true>), grid, block, 0, st, (E *)dst, value, a);
  else if (small)
    hipLaunchKernelGGL((k_fill_rows<E, EPT, int32_t, false>), grid, block, 0, st, (E *)dst, value, a);
  else
    hipLaunchKernelGGL((k_fill_rows<E, EPT, int64_t, false>), grid, block, 0, st, (E *)dst, value, a);
  return ph::check_launch("k_fill_rows");
}

int dispatch_copy(void *dst, const ph_view_desc *dd, const void *src, const ph_view_desc *sd,
                  const void *value) {
  switch (ph::dtype_size(dd->dtype)) {
    case 1: return launch_copy<uint8_t>(dst, dd, src, sd, value);
    case 2: return launch_copy<uint16_t>(dst, dd, src, sd, value);
    case 4: return launch_copy<uint32_t>(dst, dd, src, sd, value);
    case 8: return launch_copy<uint64_t>(dst, dd, src, sd, value);
  }
  return ph::set_error(PH_E_ARGUMENT, "unsupported dtype %d", dd->dtype);
}

// ---------------------------------------------------------------------------
// Fused fill + scatter over a whole row-major destination buffer.
// ---------------------------------------------------------------------------
struct FillScatterArgs {
  int32_t rank;        // destination rank
  int32_t last_axis;   // view axis of the innermost destination axis, -1 fixed
  int64_t cols;        // innermost destination extent
  int64_t rows;        // product of the outer destination extents
  int64_t ntiles, nblocks;
  int64_t shape[PH_MAX_RANK];
  int64_t first[PH_MAX_RANK];
  int64_t step[PH_MAX_RANK];
  int64_t count[PH_MAX_RANK];
  int32_t axis[PH_MAX_RANK];
  int64_t src_stride[PH_MAX_RANK];  // per view axis
  int64_t src_off;
  // innermost axis: member iff (c - first) = q*step, 0 <= q < count
  int64_t in_first, in_step, in_count, in_src_stride;
  // |in_step| as a 32-bit magic divisor: q = (umulhi(u, m) + u) >> l
  uint32_t in_d, in_m, in_l, pad2_;
};

// Granlund-Montgomery round-up divisor for 32-bit unsigned n < 2^32:
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1.
void magic_u32(uint32_t d, uint32_t *m, uint32_t *l) {
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  *l = s;
  *m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
}

__device__ __forceinline__ uint32_t fast_div(uint32_t n, uint32_t m, uint32_t l) {
  return (uint32_t)(((uint64_t)__umulhi(n, m) + n) >> l);
}

// Row membership (uniform per workgroup): decode the outer coordinate of
// destination row `row`; member rows also give the source offset of the row.
__device__ __forceinline__ bool fs_row(const FillScatterArgs &a, int64_t row, int64_t *soff_out) {
  bool member = true;
  int64_t soff = a.src_off;
  int64_t r = row;
  for (int k = a.rank - 2; k >= 0; --k) {
    const int64_t n = a.shape[k];
    const int64_t q = r / n;
    const int64_t x = r - q * n;
    r = q;
    const int64_t rel = x - a.first[k];
    const int64_t st = a.step[k];
    if (a.axis[k] < 0) {
      member = member && (rel == 0);
    } else {
      const int64_t j = rel / st;
      member = member && (j * st == rel) && j >= 0 && j < a.count[k];
      soff += j * a.src_stride[a.axis[k]];
    }
  }
  *soff_out = soff;
  return member;
}

// Vector path: a lane owns 16 consecutive bytes of a destination row (VEC
// elements) per group, G groups per thread; every store is a 16-byte
// non-temporal store (the destination is written once and not re-read here).
// Innermost-axis membership uses 32-bit arithmetic with the magic divisor.
template <typename E, int G>
__global__ __launch_bounds__(kBlock) void k_fill_scatter_vec(E *__restrict__ dst,
                                                             const E *__restrict__ src,
                                                             FillScatterArgs a, E fill) {
  constexpr int VEC = 16 / (int)sizeof(E);
  typedef E V __attribute__((ext_vector_type(VEC)));
  constexpr int64_t kTile = (int64_t)kBlock * VEC * G;
  const int32_t cols = (int32_t)a.cols;
  const int32_t in_first = (int32_t)a.in_first;
  const int32_t in_step = (int32_t)a.in_step;
  const uint32_t in_count = (uint32_t)a.in_count;
  for (int64_t b = blockIdx.x; b < a.nblocks; b += gridDim.x) {
    const int64_t row = b / a.ntiles;
    const int64_t tile = b - row * a.ntiles;
    int64_t soff;
    const bool member = fs_row(a, row, &soff);
    E *drow = dst + row * a.cols;
    if (!member) {
      V f;
#pragma unroll
      for (int v = 0; v < VEC; ++v) f[v] = fill;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int32_t c = (int32_t)(tile * kTile) + (g * kBlock + (int32_t)threadIdx.x) * VEC;
        if (c < cols) __builtin_nontemporal_store(f, (V *)(drow + c));
      }
      continue;
    }
    V out[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int32_t c0 = (int32_t)(tile * kTile) + (g * kBlock + (int32_t)threadIdx.x) * VEC;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const int32_t rel = c0 + v - in_first;
        bool ok;
        uint32_t q;
        if (a.last_axis < 0) {
          ok = rel == 0;
          q = 0;
        } else {
          const uint32_t u = (uint32_t)(in_step > 0 ? rel : -rel);
          ok = in_step > 0 ? rel >= 0 : rel <= 0;
          q = fast_div(u, a.in_m, a.in_l);
          ok = ok && q * a.in_d == u && q < in_count && c0 + v < cols;
        }
        out[g][v] = ok ? src[soff + (int64_t)q * a.in_src_stride] : fill;
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int32_t c = (int32_t)(tile * kTile) + (g * kBlock + (int32_t)threadIdx.x) * VEC;
      if (c < cols) __builtin_nontemporal_store(out[g], (V *)(drow + c));
    }
  }
}

template <typename E, int EPT>
__global__ __launch_bounds__(kBlock) void k_fill_scatter(E *__restrict__ dst,
                                                         const E *__restrict__ src,
                                                         FillScatterArgs a, E fill) {
  constexpr int64_t kTile = (int64_t)kBlock * EPT;
  for (int64_t b = blockIdx.x; b < a.nblocks; b += gridDim.x) {
    const int64_t row = b / a.ntiles;
    const int64_t tile = b - row * a.ntiles;
    // decode the outer coordinate; decide row membership (uniform)
    bool member = true;
    int64_t soff = a.src_off;
    int64_t r = row;
    for (int k = a.rank - 2; k >= 0; --k) {
      const int64_t n = a.shape[k];
      const int64_t q = r / n;
      const int64_t x = r - q * n;
      r = q;
      const int64_t rel = x - a.first[k];
      const int64_t st = a.step[k];
      if (a.axis[k] < 0) {
        member = member && (rel == 0);
      } else {
        const int64_t j = rel / st;
        member = member && (j * st == rel) && j >= 0 && j < a.count[k];
        soff += j * a.src_stride[a.axis[k]];
      }
    }
    const int64_t dbase = row * a.cols;
    const int64_t c0 = tile * kTile + threadIdx.x;
    if (!member) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int64_t c = c0 + (int64_t)e * kBlock;
        if (c < a.cols) dst[dbase + c] = fill;
      }
      continue;
    }
    E v[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int64_t c = c0 + (int64_t)e * kBlock;
      v[e] = fill;
      if (c < a.cols) {
        const int64_t rel = c - a.in_first;
        if (a.last_axis < 0) {
          if (rel == 0) v[e] = src[soff];
        } else {
          const int64_t j = rel / a.in_step;
          if (j * a.in_step == rel && j >= 0 && j < a.in_count) v[e] = src[soff + j * a.in_src_stride];
        }
      }
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int64_t c = c0 + (int64_t)e * kBlock;
      if (c < a.cols) dst[dbase + c] = v[e];
    }
  }
}

template <typename E>
int launch_fill_scatter(void *dst, FillScatterArgs a, const void *src, const void *fill_bits) {
  E fill{};
  std::memcpy(&fill, fill_bits, sizeof(E));
  constexpr int VEC = 16 / (int)sizeof(E);
  constexpr int G = 4;
  magic_u32((uint32_t)(a.in_step < 0 ? -a.in_step : a.in_step), &a.in_m, &a.in_l);
  a.in_d = (uint32_t)(a.in_step < 0 ? -a.in_step : a.in_step);
  const bool vec = a.cols % VEC == 0 && ((uintptr_t)dst % 16) == 0 && a.cols < (int64_t(1) << 30) &&
                   a.in_first < (int64_t(1) << 30) && (a.in_step < (int64_t(1) << 30)) &&
                   (a.in_step > -(int64_t(1) << 30));
  if (vec) {
    const int64_t tile = (int64_t)kBlock * VEC * G;
    a.ntiles = (a.cols + tile - 1) / tile;
    a.nblocks = a.rows * a.ntiles;
    if (a.nblocks == 0) return PH_OK;
    int rc = ph::launch_gate("fill scatter");
    if (rc) return rc;
    hipLaunchKernelGGL((k_fill_scatter_vec<E, G>), dim3((unsigned)grid_for(a.nblocks)), dim3(kBlock),
                       0, ph::stream(), (E *)dst, (const E *)src, a, fill);
    return ph::check_launch("k_fill_scatter_vec");
  }
  constexpr int EPT = sizeof(E) >= 8 ? 4 : 8;
  const int64_t tile = (int64_t)kBlock * EPT;
  a.ntiles = (a.cols + tile - 1) / tile;
  a.nblocks = a.rows * a.ntiles;
  if (a.nblocks == 0) return PH_OK;
  int rc = ph::launch_gate("fill scatter");
  if (rc) return rc;
  hipLaunchKernelGGL((k_fill_scatter<E, EPT>), dim3((unsigned)grid_for(a.nblocks)), dim3(kBlock),
                     0, ph::stream(), (E *)dst, (const E *)src, a, fill);
  return ph::check_launch("k_fill_scatter");
}

// ---------------------------------------------------------------------------
// Seeded random fill (counter-based splitmix64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t draw(uint64_t seed, uint64_t k) {
  return mix64(seed + (k + 1) * 0x9E3779B97F4A7C15ULL);
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(kBlock) void k_random(void *buf, int dtype, int64_t n, uint64_t seed,
                                                   uint64_t offset, int dist, double a, double b) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint64_t k = offset + (uint64_t)i;
    double x = 0.0;
    int64_t xi = 0;
    if (dist == 0) {
      const uint64_t d = draw(seed, k);
      if (dtype == PH_I32) {
        const int64_t span = (int64_t)(b - a);
        xi = (int64_t)a + (int64_t)((d >> 11) % (uint64_t)span);
      } else {
        const double u = (double)(d >> 11) * 0x1.0p-53;
        x = __dadd_rn(a, __dmul_rn(__dsub_rn(b, a), u));
      }
    } else if (dist == 1) {
      const uint64_t kp = k & ~1ULL;
      const double u1 = (double)(draw(seed, kp) >> 11) * 0x1.0p-53;
      const double u2 = (double)(draw(seed, kp + 1) >> 11) * 0x1.0p-53;
      const double r = sqrt(-2.0 * log(1.0 - u1));
      const double t = __dmul_rn(two_pi, u2);
      const double z = (k & 1) ? __dmul_rn(r, sin(t)) : __dmul_rn(r, cos(t));
      x = __dadd_rn(a, __dmul_rn(b, z));
    } else {
      xi = (int64_t)(draw(seed, k) >> 33);
    }
    switch (dtype) {
      case PH_I32: ((int32_t *)buf)[i] = (int32_t)xi; break;
      case PH_F32: ((float *)buf)[i] = (float)x; break;
      case PH_F64: ((double *)buf)[i] = x; break;
      case PH_BF16: ((uint16_t *)buf)[i] = f32_to_bf16_rne((float)x); break;
    }
  }
}

__global__ void k_clear_i32(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

int ph::clear_i32(int *p, int n) {
  if (n <= 0) return PH_OK;
  hipLaunchKernelGGL(k_clear_i32, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, ph::stream(), p, n);
  return ph::check_launch("k_clear_i32");
}

extern "C" {

int ph_fill_random(void *buf, int32_t dtype, int64_t n, uint64_t seed, uint64_t offset,
                   int32_t dist, double a, double b) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (n <= 0) return PH_OK;
  if (dtype == PH_BOOL || ph::dtype_size(dtype) == 0)
    return ph::set_error(PH_E_ARGUMENT, "ph_fill_random: unsupported dtype %d", dtype);
  if (dist < 0 || dist > 2) return ph::set_error(PH_E_ARGUMENT, "ph_fill_random: dist %d", dist);
  if (dist == 0 && dtype == PH_I32 && !(b > a))
    return ph::set_error(PH_E_ARGUMENT, "ph_fill_random: empty integer range");
  int rc = ph::launch_gate("random");
  if (rc) return rc;
  hipLaunchKernelGGL(k_random, dim3((unsigned)grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, ph::stream(), buf, dtype, n, seed, offset, dist, a, b);
  return ph::check_launch("k_random");
}

int ph_fill(void *buf, const ph_view_desc *dst, const void *value) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  int rc = ph::validate_desc(dst, "fill destination");
  if (rc) return rc;
  if (!value) return ph::set_error(PH_E_ARGUMENT, "ph_fill: null value");
  return dispatch_copy(buf, dst, nullptr, nullptr, value);
}

int ph_strided_copy(void *dst, const ph_view_desc *dd, const void *src, const ph_view_desc *sd,
                    int32_t flags) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  int rc = ph::validate_desc(dd, "copy destination");
  if (rc) return rc;
  rc = ph::validate_desc(sd, "copy source");
  if (rc) return rc;
  if (dd->dtype != sd->dtype && ph::dtype_size(dd->dtype) != ph::dtype_size(sd->dtype))
    return ph::set_error(PH_E_ARGUMENT, "copy between dtypes %d and %d", sd->dtype, dd->dtype);
  if (dd->view.rank != sd->view.rank)
    return ph::set_error(PH_E_SHAPE, "copy between ranks %d and %d", sd->view.rank, dd->view.rank);
  for (int i = 0; i < dd->view.rank; ++i)
    if (dd->view.shape[i] != sd->view.shape[i])
      return ph::set_error(PH_E_SHAPE, "copy shapes differ on axis %d (%lld vs %lld)", i,
                           (long long)sd->view.shape[i], (long long)dd->view.shape[i]);
  if (flags != 0) return ph::set_error(PH_E_ARGUMENT, "ph_strided_copy: unknown flags %d", flags);
  if (ph::desc_count(dd) == 0) return PH_OK;
  const size_t es = ph::dtype_size(dd->dtype);
  if (src == dst) {
    // the reference's order decides what a source reading the destination
    // sees; the parallel kernels below are used when it cannot matter
    bool done = false;
    rc = ph::ordered_copy(dst, dd, src, sd, &done);
    if (rc || done) return rc;
  } else {
    const char *d0 = (const char *)dst, *s0 = (const char *)src;
    if (!(s0 + (size_t)sd->extent * es <= d0 || d0 + (size_t)dd->extent * es <= s0))
      return ph::set_error(PH_E_ARGUMENT, "copy source overlaps the destination through another base pointer");
  }
  return dispatch_copy(dst, dd, src, sd, nullptr);
}

int ph_view_index(int64_t *out, const ph_view_desc *d) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  int rc = ph::validate_desc(d, "view_index");
  if (rc) return rc;
  const int64_t total = ph::desc_count(d);
  if (total == 0) return PH_OK;
  if (!out) return ph::set_error(PH_E_ARGUMENT, "ph_view_index: null output");
  rc = ph::launch_gate("view_index");
  if (rc) return rc;
  hipLaunchKernelGGL(k_view_index, dim3((unsigned)grid_for((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     ph::stream(), out, *d, total);
  return ph::check_launch("k_view_index");
}

int ph_fill_scatter(void *dst, int32_t dtype, const int64_t *dst_shape, int32_t rank,
                    const int64_t *first, const int64_t *step, const int64_t *count,
                    const int32_t *axis, const void *src, const ph_view_desc *sd,
                    const void *fill) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (rank < 1 || rank > PH_MAX_RANK)
    return ph::set_error(PH_E_DIMENSION, "ph_fill_scatter: rank %d", rank);
  int rc = ph::validate_desc(sd, "fill_scatter source");
  if (rc) return rc;
  if (sd->nstage != 0)
    return ph::set_error(PH_E_UNSUPPORTED, "ph_fill_scatter: source must be affine");
  if (ph::dtype_size(dtype) != ph::dtype_size(sd->dtype))
    return ph::set_error(PH_E_ARGUMENT, "ph_fill_scatter: dtype mismatch");
  FillScatterArgs a;
  std::memset(&a, 0, sizeof(a));
  a.rank = rank;
  a.rows = 1;
  int used[PH_MAX_RANK] = {0};
  for (int k = 0; k < rank; ++k) {
    if (dst_shape[k] < 0) return ph::set_error(PH_E_DIMENSION, "negative extent");
    a.shape[k] = dst_shape[k];
    a.first[k] = first[k];
    a.step[k] = step[k];
    a.count[k] = count[k];
    a.axis[k] = axis[k];
    if (axis[k] >= 0) {
      if (axis[k] >= sd->view.rank || used[axis[k]])
        return ph::set_error(PH_E_ARGUMENT, "ph_fill_scatter: bad view axis %d", axis[k]);
      used[axis[k]] = 1;
      if (count[k] != sd->view.shape[axis[k]])
        return ph::set_error(PH_E_SHAPE, "ph_fill_scatter: axis %d count %lld vs source %lld", k,
                             (long long)count[k], (long long)sd->view.shape[axis[k]]);
      if (count[k] > 0) {
        if (step[k] == 0) return ph::set_error(PH_E_ARGUMENT, "ph_fill_scatter: zero step");
        const int64_t lastc = first[k] + (count[k] - 1) * step[k];
        if (first[k] < 0 || first[k] >= dst_shape[k] || lastc < 0 || lastc >= dst_shape[k])
          return ph::set_error(PH_E_INDEX, "ph_fill_scatter: region leaves axis %d", k);
      }
    } else if (first[k] < 0 || first[k] >= dst_shape[k]) {
      if (dst_shape[k] > 0) return ph::set_error(PH_E_INDEX, "ph_fill_scatter: fixed index outside axis %d", k);
    }
    if (k < rank - 1) a.rows *= dst_shape[k];
  }
  for (int i = 0; i < sd->view.rank; ++i) {
    if (!used[i] && sd->view.shape[i] != 1)
      return ph::set_error(PH_E_SHAPE, "ph_fill_scatter: source axis %d not mapped", i);
    a.src_stride[i] = sd->view.stride[i];
  }
  a.src_off = sd->view.offset;
  a.cols = dst_shape[rank - 1];
  a.last_axis = axis[rank - 1];
  a.in_first = first[rank - 1];
  a.in_step = step[rank - 1] == 0 ? 1 : step[rank - 1];
  a.in_count = count[rank - 1];
  a.in_src_stride = axis[rank - 1] >= 0 ? sd->view.stride[axis[rank - 1]] : 0;
  switch (ph::dtype_size(dtype)) {
    case 1: return launch_fill_scatter<uint8_t>(dst, a, src, fill);
    case 2: return launch_fill_scatter<uint16_t>(dst, a, src, fill);
    case 4: return launch_fill_scatter<uint32_t>(dst, a, src, fill);
    case 8: return launch_fill_scatter<uint64_t>(dst, a, src, fill);
  }
  return ph::set_error(PH_E_ARGUMENT, "ph_fill_scatter: dtype %d", dtype);
}

}  // extern "C"
