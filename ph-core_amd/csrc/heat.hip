// Explicit 2-D heat equation (5-point FTCS, Dirichlet zero outside the grid)
// and the FP64 reductions around it (total energy, relative L2 distance).
//
// Step kernel (k_heat_multi): K time steps per pass over the grid.  A
// wavefront owns a strip of columns and sweeps a band of rows top to bottom,
// keeping two rows per time level in registers, so each pass reads the grid
// once (plus 2K halo rows per band and two or three overlapping lanes per
// strip edge) and writes it once.  Loads and stores are VEC-wide (16 or 32 bytes per
// lane) and coalesced.  Workgroups are numbered so that strips adjacent in
// the grid run on the same XCD (workgroups are dealt to the 8 XCDs
// round-robin), which keeps the overlapping columns in that XCD's L2.
//
// The cell update keeps the operation order of examples/heat_equation.cr:47
// and :50 extended to two axes:
//   d = (((uN - 4u) + uS) + uW) + uE;  u' = u + d * r
// with one contraction that changes no result: 4u is exact unless it
// overflows, so uN - 4u rounds once whether it is computed as a multiply and
// a subtract or as one fused multiply-add.  The fused form runs only when the
// host has shown that 4u cannot overflow in the passes ahead (fused_safe).
// Float32 grids (the build's own variant of the workload; the reference's
// example is Float64) also compute u + d * r as one fused multiply-add,
// which rounds once instead of twice: one instruction less per cell and
// step, with the accuracy stated against the Float64 run.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>
#include <type_traits>

#include "common.hpp"

namespace {

constexpr int kWaves = 4;    // wavefronts (strips) per workgroup

template <typename T> struct Vec;
template <> struct Vec<float> { typedef float4 type; static constexpr int n = 4; };
template <> struct Vec<double> { typedef double4 type; static constexpr int n = 4; };

template <typename T>
__device__ __forceinline__ void unpack(const typename Vec<T>::type &v, T *o) {
  if constexpr (Vec<T>::n == 4) {
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else {
    o[0] = v.x; o[1] = v.y;
  }
}
template <typename T>
__device__ __forceinline__ typename Vec<T>::type pack(const T *o) {
  typename Vec<T>::type v;
  if constexpr (Vec<T>::n == 4) {
    v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
  } else {
    v.x = o[0]; v.y = o[1];
  }
  return v;
}

// Fused multiply-add in the cell's own precision.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Whole-wavefront lane shifts by one (DPP wave_shr:1 / wave_shl:1): lane i
// receives lane i-1's (shr) or lane i+1's (shl) value; the lane at the end
// receives 0.  One VALU move per 32 bits instead of an LDS-crossbar
// permute.  Only lanes 0 and 63 see the 0, and the stencil never stores them.
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;
template <int kCtrl>
__device__ __forceinline__ float lane_shift(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, true));
}
template <int kCtrl>
__device__ __forceinline__ double lane_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), kCtrl, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), kCtrl, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

struct HeatArgs {
  int64_t ny, nx;
  int64_t ty;        // rows per band
  int64_t nstrips;   // strips per row band
  int64_t nbands;
  int64_t nblocks;   // workgroups
  int64_t per_xcd;   // workgroups per XCD slice
};

// Calls f(std::integral_constant<int, G>) for G = 0 .. N-1.
template <typename F, int... G>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, G...>) {
  (f(std::integral_constant<int, G>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// K time steps per pass (temporal blocking, 1 <= K <= kHalo * VEC).
//
// Columns: strips overlap.  A wavefront holds 64*VEC consecutive columns and
// writes the (64 - 2*kHalo)*VEC in the middle (lanes kHalo .. 63-kHalo).
// West/east neighbours come from the adjacent lane (a DPP lane shift, which
// hands the lanes at the wavefront's two ends a zero); those edge lanes have
// no neighbour beyond the wavefront, so their values go wrong by one column
// per level, and after K <= kHalo*VEC levels only the kHalo lanes at each
// edge are wrong: they are not stored.  When rows allow whole-vector lanes
// and hold at least one wavefront's columns (the usual case), the first and
// last strips are placed flush with the grid's edges: there the zero of the
// lane shift is exactly the zero outside the grid, so those edge lanes are
// exact and stored, and no lane lies outside the grid.
// Rows: level l+1 of row y is computed from level-l rows y-1, y, y+1.  The
// wavefront keeps, per level, the last two rows in registers; each loaded
// level-0 row y+K yields level-1 row y+K-1, ..., level-K row y.  A band's
// sweep starts 2K rows above it, so every level is exact before the band's
// first output row (level l is exact 2l rows after the start).  The band at
// the top of the grid starts at the grid's edge instead, with every level of
// the rows above the edge held at zero (the boundary value; for its first K
// rows by masks fixed at compile time), and the band at the bottom does the
// same sweeping upwards (north and south trade places in the registers, not
// in the formula).  Bands that fit neither case, and grids whose rows do not
// allow whole-vector lanes, take a sweep that masks every value outside the
// grid.  Each value is computed with the single-step formula, so K steps per
// pass are bit-identical to K single steps.
// ---------------------------------------------------------------------------
template <typename T, int K, int kAhead, int kHalo, bool kFused>
__global__ __launch_bounds__(64 * kWaves) void k_heat_multi(const T *__restrict__ u, T *__restrict__ out,
                                                            HeatArgs a, T r, int vec_ok) {
  constexpr int V = Vec<T>::n;
  static_assert(K >= 1 && K <= kHalo * V, "K steps per pass need K <= kHalo * VEC halo columns");
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t b = blockIdx.x;
  const int64_t lb = (b % 8) * a.per_xcd + b / 8;
  if (lb >= a.nblocks) return;
  const int64_t gpb = (a.nstrips + kWaves - 1) / kWaves;
  const int64_t strip = (lb % gpb) * kWaves + wave;
  const int64_t band = lb / gpb;
  if (strip >= a.nstrips) return;
  const int64_t y0 = band * a.ty;
  const int64_t y1 = (y0 + a.ty < a.ny) ? y0 + a.ty : a.ny;
  const T four = T(4);
  const bool flush = vec_ok && a.nx >= 64 * V;
  int64_t xs = strip * (64 - 2 * kHalo) * V - kHalo * V;  // the wavefront's first column
  if (flush) xs = xs < 0 ? 0 : (xs > a.nx - 64 * V ? a.nx - 64 * V : xs);
  const int64_t x = xs + (int64_t)lane * V;  // this lane's first column
  const bool west_edge = flush && xs == 0, east_edge = flush && xs + 64 * V == a.nx;
  const bool store_lane = (lane >= kHalo || west_edge) && (lane < 64 - kHalo || east_edge);

  // One level update of V cells: north n, centre c (with the lane's west and
  // east neighbour columns wl, er), south s.
  auto level = [&](const T (&n)[V], const T (&c)[V], const T (&s)[V], T wl, T er, T (&o)[V]) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const T w = i == 0 ? wl : c[i - 1];
      const T e = i == V - 1 ? er : c[i + 1];
      T nc;
      if constexpr (kFused) nc = fma_t(-four, c[i], n[i]);  // exact 4u: one rounding either way
      else nc = n[i] - four * c[i];
      const T d = ((nc + s[i]) + w) + e;
      if constexpr (sizeof(T) == 4) o[i] = fma_t(d, r, c[i]);  // Float32: u + d*r rounded once
      else o[i] = c[i] + d * r;
    }
  };

  // ---- Sweep with no masks (flush strips).  Rows are counted in the
  // sweep's direction, m = y downwards or m = ny-1-y upwards; the band is
  // [m0, m1).  kEdge: the band starts at the grid's edge (m0 = 0).
  auto clean = [&](auto up_tag, auto edge_tag, int64_t m0, int64_t m1) {
    constexpr bool kUp = decltype(up_tag)::value;
    constexpr bool kEdge = decltype(edge_tag)::value;
    auto row_at = [&](auto *base, int64_t m) { return base + (kUp ? a.ny - 1 - m : m) * a.nx + x; };
    T prev[K][2][V];  // per level l: rows m+K-l-2 and m+K-l-1 (in sweep order)
#pragma unroll
    for (int l = 0; l < K; ++l)
#pragma unroll
      for (int i = 0; i < V; ++i) prev[l][0][i] = prev[l][1][i] = T(0);
    // Level-0 rows come through a ring of kAhead slots, each refilled right
    // after it is consumed, so kAhead row loads stay in flight.  Rows past
    // m1 + K - 1 are never consumed; they read that row again, so every load
    // is issued unconditionally (no branch separates it from the wait for
    // the loads before it).
    auto load_level0 = [&](int64_t m, T *dst) {
      const int64_t mc = m < m1 + K ? m : m1 + K - 1;
      unpack<T>(*(const typename Vec<T>::type *)row_at(u, mc), dst);
    };
    const int64_t start = m0 - (kEdge ? K : 2 * K);
    T L[kAhead][V];  // slot k holds level-0 row m+k+K
#pragma unroll
    for (int k = 0; k < kAhead; ++k) load_level0(start + k + K, L[k]);
    // kAhead rows from m = y.  J0 (compile time): in an edge sweep, the
    // index of the group's first row among the first K rows; K otherwise.
    auto group = [&](int64_t y, auto j0_tag) {
      constexpr int J0 = decltype(j0_tag)::value;
#pragma unroll
      for (int k = 0; k < kAhead; ++k) {
        // the last group may run past m1: those rows are computed from values
        // that are never stored, and not stored themselves
        const int64_t mm = y + k;
        T cur[V];
#pragma unroll
        for (int i = 0; i < V; ++i) cur[i] = L[k][i];
        load_level0(mm + kAhead + K, L[k]);  // refill the slot kAhead rows ahead
#pragma unroll
        for (int l = 0; l < K; ++l) {
          // level l+1 of row mm+K-l-1 from prev[l][0], prev[l][1], cur.  In
          // row j = J0 + k of an edge sweep, levels l >= j belong to rows
          // beyond the edge: they stay 0.
          T nxt[V];
          if (J0 + k < K && l >= J0 + k) {
#pragma unroll
            for (int i = 0; i < V; ++i) nxt[i] = T(0);
          } else {
            const T wl = lane_shift<kDppWaveShr1>(prev[l][1][V - 1]);  // lane-1's last column
            const T er = lane_shift<kDppWaveShl1>(prev[l][1][0]);      // lane+1's first column
            if constexpr (kUp) level(cur, prev[l][1], prev[l][0], wl, er, nxt);
            else level(prev[l][0], prev[l][1], cur, wl, er, nxt);
          }
#pragma unroll
          for (int i = 0; i < V; ++i) {
            prev[l][0][i] = prev[l][1][i];
            prev[l][1][i] = cur[i];
            cur[i] = nxt[i];
          }
        }
        // cur: level K of row mm
        if (mm >= m0 && mm < m1 && store_lane)
          *(typename Vec<T>::type *)row_at(out, mm) = pack<T>(cur);
      }
    };
    int64_t y = start;
    if constexpr (kEdge) {
      constexpr int kGroups = (K + kAhead - 1) / kAhead;
      static_for<kGroups>([&](auto g) {
        constexpr int G = decltype(g)::value;
        group(y + G * kAhead, std::integral_constant<int, G * kAhead>{});
      });
      y += kGroups * kAhead;
    }
    for (; y < m1; y += kAhead) group(y, std::integral_constant<int, K>{});
  };

  if (flush) {
    if (y0 - 2 * K >= 0 && y1 + K <= a.ny) {
      clean(std::false_type{}, std::false_type{}, y0, y1);
      return;
    }
    if (y0 == 0 && y1 + K <= a.ny) {
      clean(std::false_type{}, std::true_type{}, 0, y1);
      return;
    }
    if (y1 == a.ny && y0 >= K) {
      clean(std::true_type{}, std::true_type{}, 0, a.ny - y0);
      return;
    }
  }

  // ---- Sweep with masks: every value outside the grid is held at zero.
  const bool full = vec_ok && x >= 0 && x + V <= a.nx;  // whole aligned vector inside the row
  bool colok[V];
#pragma unroll
  for (int i = 0; i < V; ++i) colok[i] = x + i >= 0 && x + i < a.nx;
  T prev[K][2][V];
#pragma unroll
  for (int l = 0; l < K; ++l)
#pragma unroll
    for (int i = 0; i < V; ++i) prev[l][0][i] = prev[l][1][i] = T(0);
  auto load_level0 = [&](int64_t ry, T *dst) {
    if (ry >= 0 && ry < a.ny && ry < y1 + K) {
      const T *row = u + ry * a.nx;
      if (full) {
        unpack<T>(*(const typename Vec<T>::type *)(row + x), dst);
      } else {
#pragma unroll
        for (int i = 0; i < V; ++i) dst[i] = colok[i] ? row[x + i] : T(0);
      }
    } else {
#pragma unroll
      for (int i = 0; i < V; ++i) dst[i] = T(0);
    }
  };
  T L[kAhead][V];
#pragma unroll
  for (int k = 0; k < kAhead; ++k) load_level0(y0 - 2 * K + k + K, L[k]);
  for (int64_t y = y0 - 2 * K; y < y1; y += kAhead) {
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      const int64_t yy = y + k;
      T cur[V];
#pragma unroll
      for (int i = 0; i < V; ++i) cur[i] = L[k][i];
      load_level0(yy + kAhead + K, L[k]);
#pragma unroll
      for (int l = 0; l < K; ++l) {
        const int64_t row = yy + K - l - 1;
        const bool rowok = row >= 0 && row < a.ny;
        const T wl = lane_shift<kDppWaveShr1>(prev[l][1][V - 1]);
        const T er = lane_shift<kDppWaveShl1>(prev[l][1][0]);
        T nxt[V];
        level(prev[l][0], prev[l][1], cur, wl, er, nxt);
#pragma unroll
        for (int i = 0; i < V; ++i) {
          nxt[i] = (rowok && colok[i]) ? nxt[i] : T(0);
          prev[l][0][i] = prev[l][1][i];
          prev[l][1][i] = cur[i];
          cur[i] = nxt[i];
        }
      }
      if (yy >= y0 && yy < y1 && store_lane) {
        T *orow = out + yy * a.nx;
        if (full) {
          *(typename Vec<T>::type *)(orow + x) = pack<T>(cur);
        } else {
#pragma unroll
          for (int i = 0; i < V; ++i)
            if (colok[i]) orow[x + i] = cur[i];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Deterministic FP64 sums: pass 1 gives each workgroup a fixed contiguous
// chunk (Neumaier-compensated per thread, ordered tree per group); pass 2
// folds the per-group partials in order in one workgroup.
// ---------------------------------------------------------------------------
struct D2 {
  double s, c;
};
__device__ __forceinline__ void d2_add(D2 &a, double v) {
  const double t = a.s + v;
  const double d = (fabs(a.s) >= fabs(v)) ? ((a.s - t) + v) : ((v - t) + a.s);
  a.c += d;
  a.s = t;
}
__device__ __forceinline__ D2 d2_comb(const D2 &a, const D2 &b) {
  const double t = a.s + b.s;
  const double bb = t - a.s;
  const double err = (a.s - (t - bb)) + (b.s - bb);
  D2 r;
  r.s = t;
  r.c = (a.c + b.c) + err;
  return r;
}

template <int K>
__device__ __forceinline__ void block_fold(D2 (&v)[K], D2 (*sh)[256]) {
  for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    D2 o[K];
    for (int k = 0; k < K; ++k) {
      o[k] = sh[k][threadIdx.x];
      if ((threadIdx.x & (2 * off - 1)) == 0) o[k] = d2_comb(o[k], sh[k][threadIdx.x + off]);
    }
    __syncthreads();
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = o[k];
    __syncthreads();
  }
  for (int k = 0; k < K; ++k) v[k] = sh[k][0];
}

// mode 0: K=1 sum of a; mode 1: K=2 sums of (a-b)^2 and b^2
template <typename TA, typename TB, int K>
__global__ __launch_bounds__(256) void k_sum_pass1(const TA *__restrict__ a,
                                                   const TB *__restrict__ b, int64_t n,
                                                   int64_t chunk, D2 *part) {
  __shared__ D2 sh[K][256];
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = (lo + chunk < n) ? lo + chunk : n;
  D2 acc[K];
  for (int k = 0; k < K; ++k) acc[k] = D2{0.0, 0.0};
  // lane-strided (coalesced); the order is fixed, so the result is
  // deterministic
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
    const double x = (double)a[i];
    if constexpr (K == 1) {
      d2_add(acc[0], x);
    } else {
      const double y = (double)b[i];
      const double d = x - y;
      d2_add(acc[0], d * d);
      d2_add(acc[1], y * y);
    }
  }
  block_fold<K>(acc, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) part[(int64_t)blockIdx.x * K + k] = acc[k];
}

template <int K>
__global__ __launch_bounds__(256) void k_sum_pass2(const D2 *__restrict__ part, int64_t nparts,
                                                   double *out) {
  __shared__ D2 sh[K][256];
  const int64_t per = (nparts + 255) / 256;
  const int64_t lo = (int64_t)threadIdx.x * per;
  const int64_t hi = (lo + per < nparts) ? lo + per : nparts;
  D2 acc[K];
  for (int k = 0; k < K; ++k) acc[k] = D2{0.0, 0.0};
  for (int64_t i = lo; i < hi; ++i)
    for (int k = 0; k < K; ++k) acc[k] = d2_comb(acc[k], part[i * K + k]);
  block_fold<K>(acc, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) out[k] = acc[k].s + acc[k].c;
}

// `part_buf` (cap bytes), when given, holds the per-group partials instead
// of library scratch: at most 2048 groups, so 2048 * K * 16 bytes cover any n.
template <typename TA, typename TB, int K>
int fp64_sum(const TA *a, const TB *b, int64_t n, double *out, void *part_buf = nullptr, size_t cap = 0) {
  const int64_t nparts_max = 2048;
  int64_t chunk = (n + nparts_max - 1) / nparts_max;
  if (chunk < 4096) chunk = 4096;
  const int64_t nparts = (n + chunk - 1) / chunk;
  const size_t need = (size_t)(nparts > 0 ? nparts : 1) * K * sizeof(D2);
  void *part = part_buf;
  int rc = PH_OK;
  if (part) {
    if (need > cap)
      return ph::set_error(PH_E_ARGUMENT, "fp64 sum: %zu bytes of partials exceed the fixed %zu", need, cap);
  } else {
    rc = ph::scratch(need, &part, 2);
    if (rc) return rc;
  }
  rc = ph::launch_gate("fp64 sum");
  if (rc) return rc;
  if (nparts > 0)
    hipLaunchKernelGGL((k_sum_pass1<TA, TB, K>), dim3((unsigned)nparts), dim3(256), 0, ph::stream(),
                       a, b, n, chunk, (D2 *)part);
  hipLaunchKernelGGL((k_sum_pass2<K>), dim3(1), dim3(256), 0, ph::stream(), (const D2 *)part,
                     nparts, out);
  return ph::check_launch("k_sum");
}

// Workgroups of kernel `f` that are resident at once on the current device
// (compute units x workgroups per compute unit at the kernel's register
// use), cached per device and kernel.
int64_t resident_groups(const void *f) {
  // a fixed table (no allocation on any call): one slot per kernel and device
  struct Slot {
    int dev;
    const void *f;
    int64_t groups;
  };
  static Slot cache[128];
  static int used = 0;
  const int dev = ph::device();
  for (int k = 0; k < used; ++k)
    if (cache[k].dev == dev && cache[k].f == f) return cache[k].groups;
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, 64 * kWaves, 0) != hipSuccess || per < 1) per = 1;
  const int64_t g = (int64_t)cus * per;
  if (used < 128) cache[used++] = Slot{dev, f, g};
  return g;
}

// Band height for a pass of K steps.  Every band is a row of independent
// workgroups, and each of them sweeps its band's rows plus 2K warm-up rows;
// the workgroups run in ceil(blocks / resident) rounds of equal length, so
// the pass takes about rounds * (ty + 2K) row sweeps.  The height minimises
// that (the smallest band count among equals) over heights that keep every
// band on a sweep without masks: two bands or more, of at least 2K rows, the
// last of at least K rows.  Small grids where no height does take the
// cheapest one.
int64_t band_rows(int64_t ny, int64_t groups_per_band, int64_t resident, int k) {
  int64_t best_ty = ny, best_cost = -1;
  bool best_clean = false;
  for (int64_t nb = 1; nb <= ny; ++nb) {
    const int64_t ty = (ny + nb - 1) / nb;
    if (ty < 8 && nb > 1) break;
    const int64_t nbands = (ny + ty - 1) / ty;
    const int64_t last = ny - (nbands - 1) * ty;
    const bool clean = nbands >= 2 && ty >= 2 * k && last >= k;
    const int64_t rounds = (groups_per_band * nbands + resident - 1) / resident;
    const int64_t cost = rounds * (ty + 2 * k);
    if (best_cost < 0 || (clean && !best_clean) || (clean == best_clean && cost < best_cost)) {
      best_cost = cost;
      best_ty = ty;
      best_clean = clean;
    }
  }
  return best_ty;
}

// `a` carries the grid and strip count; the band height is `ty` when
// positive, else chosen by band_rows for this kernel's residency.
template <typename T, int K, int A, bool F>
int launch_ahead(const T *cur, T *nxt, HeatArgs a, int64_t ty, T r, int vec_ok) {
  constexpr int H = K > 8 ? 3 : 2;  // halo lanes on each side of a wavefront's strip
  const void *f = (const void *)k_heat_multi<T, K, A, H, F>;
  constexpr int64_t wcols = (64 - 2 * H) * Vec<T>::n;  // columns a wavefront writes
  a.nstrips = (a.nx + wcols - 1) / wcols;
  const int64_t groups_per_band = (a.nstrips + kWaves - 1) / kWaves;
  a.ty = ty > 0 ? ty : band_rows(a.ny, groups_per_band, resident_groups(f), K);
  a.nbands = (a.ny + a.ty - 1) / a.ty;
  a.nblocks = groups_per_band * a.nbands;
  a.per_xcd = (a.nblocks + 7) / 8;
  const unsigned grid = (unsigned)(a.per_xcd * 8);
  hipLaunchKernelGGL((k_heat_multi<T, K, A, H, F>), dim3(grid), dim3(64 * kWaves), 0, ph::stream(), cur, nxt, a, r,
                     vec_ok);
  return ph::check_launch("k_heat_multi");
}

// Rows loaded ahead: 8 for passes of one or two steps (little arithmetic per
// row, so more loads must be in flight to keep HBM busy), else 3; `ahead`
// (PH_HEAT_AHEAD) picks 2, 3 or, for such short passes, 8.
template <typename T, int K, bool F>
int launch_multi(const T *cur, T *nxt, HeatArgs a, int64_t ty, T r, int vec_ok, int ahead) {
  if (ahead <= 0) ahead = K <= 2 ? 8 : 3;
  if constexpr (K <= 2) {
    if (ahead >= 8) return launch_ahead<T, K, 8, F>(cur, nxt, a, ty, r, vec_ok);
  }
  if (ahead >= 3) return launch_ahead<T, K, 3, F>(cur, nxt, a, ty, r, vec_ok);
  return launch_ahead<T, K, 2, F>(cur, nxt, a, ty, r, vec_ok);
}

template <typename T, bool F>
int launch_kf(int k, const T *cur, T *nxt, const HeatArgs &a, int64_t ty, T r, int vec_ok, int ahead) {
  if constexpr (sizeof(T) == 4) {  // Float32 also runs passes of twelve steps (three halo lanes)
    if (k == 12) return launch_multi<T, 12, F>(cur, nxt, a, ty, r, vec_ok, ahead);
  }
  switch (k) {
    case 8: return launch_multi<T, 8, F>(cur, nxt, a, ty, r, vec_ok, ahead);
    case 7: return launch_multi<T, 7, F>(cur, nxt, a, ty, r, vec_ok, ahead);
    case 6: return launch_multi<T, 6, F>(cur, nxt, a, ty, r, vec_ok, ahead);
    case 5: return launch_multi<T, 5, F>(cur, nxt, a, ty, r, vec_ok, ahead);
    case 4: return launch_multi<T, 4, F>(cur, nxt, a, ty, r, vec_ok, ahead);
    case 3: return launch_multi<T, 3, F>(cur, nxt, a, ty, r, vec_ok, ahead);
    case 2: return launch_multi<T, 2, F>(cur, nxt, a, ty, r, vec_ok, ahead);
    default: return launch_multi<T, 1, F>(cur, nxt, a, ty, r, vec_ok, ahead);
  }
}

template <typename T>
int launch_k(int k, bool fused, const T *cur, T *nxt, const HeatArgs &a, int64_t ty, T r, int vec_ok, int ahead) {
  return fused ? launch_kf<T, true>(k, cur, nxt, a, ty, r, vec_ok, ahead)
               : launch_kf<T, false>(k, cur, nxt, a, ty, r, vec_ok, ahead);
}

// Largest |u| of the grid as an ordered bit pattern: with the sign bit
// cleared, finite values order as their magnitudes, and infinities and NaN
// order above every finite value.
template <typename T> struct Bits;
template <> struct Bits<float> {
  typedef unsigned int U;
  static constexpr U kMag = 0x7fffffffu;
  static constexpr U kLimit = 251u << 23;  // 2^124
  static constexpr double kEps = 0x1p-24;
};
template <> struct Bits<double> {
  typedef unsigned long long U;
  static constexpr U kMag = 0x7fffffffffffffffull;
  static constexpr U kLimit = 2043ull << 52;  // 2^1020
  static constexpr double kEps = 0x1p-53;
};

template <typename T>
__global__ __launch_bounds__(256) void k_absmax_bits(const T *__restrict__ u, int64_t n, typename Bits<T>::U *out) {
  using U = typename Bits<T>::U;
  const U *b = reinterpret_cast<const U *>(u);
  U m = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const U v = b[i] & Bits<T>::kMag;
    m = v > m ? v : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const U o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// Whether the next passes may compute uN - 4u as one fused multiply-add:
// 4u must not overflow in any of them.  For 0 <= r <= 1/4 the update is a
// convex combination of a cell and its neighbours (weights 1 - 4r and r),
// so the largest |u| cannot grow beyond rounding: each step's roundings add
// at most 11 eps of it (d carries four roundings of terms summing to at most
// 8 max|u|, and the product and the final sum one each).  Starting from
// max|u| <= 2^124 (Float32) or 2^1020 (Float64), 4u stays finite for
// ln 4 / (11 eps) steps (about 2.1 million for Float32); *steps_ok returns
// a bound within that, after which the host checks again.  A grid holding
// an infinity, a NaN or a larger value, or r outside [0, 1/4], keeps the
// separate multiply and subtract.
template <typename T>
int fused_safe(const T *u, int64_t n, T r, bool *safe, int64_t *steps_ok) {
  using U = typename Bits<T>::U;
  *safe = false;
  *steps_ok = 0;
  if (!(r >= T(0) && r <= T(0.25))) return PH_OK;
  void *p;
  int rc = ph::scratch(64, &p, 2);
  if (rc) return rc;
  rc = ph::clear_i32((int *)p, 2);
  if (rc) return rc;
  rc = ph::launch_gate("heat max|u|");
  if (rc) return rc;
  const int64_t blocks = std::min<int64_t>(2048, (n + 255) / 256);
  hipLaunchKernelGGL(k_absmax_bits<T>, dim3((unsigned)std::max<int64_t>(1, blocks)), dim3(256), 0, ph::stream(), u,
                     n, (U *)p);
  rc = ph::check_launch("k_absmax_bits");
  if (rc) return rc;
  U m = 0;
  PH_HIP(hipMemcpyAsync(&m, p, sizeof(U), hipMemcpyDeviceToHost, ph::stream()));
  PH_HIP(hipStreamSynchronize(ph::stream()));
  *safe = m <= Bits<T>::kLimit;
  *steps_ok = (int64_t)(0.5 * std::log(4.0) / (11.0 * Bits<T>::kEps));
  return PH_OK;
}

// part_buf / part_cap: a fixed buffer for the energy partials (the
// allocation-free entry point); allow_fused: the fused first operation may
// be used (it needs one read-back of max|u|).
template <typename T>
int heat_run(T *u, T *tmp, int64_t ny, int64_t nx, double r, int64_t steps, int64_t every,
             double *energy, void *part_buf = nullptr, size_t part_cap = 0, bool allow_fused = true) {
  constexpr int V = Vec<T>::n;
  // Steps per pass, measured on MI355X at 8192^2: 12 for Float32 (three
  // halo lanes; 8, 10, 14 and 16 are slower) and 8 for Float64 (its 252
  // registers leave no room for more).  The rows loaded ahead follow from
  // the steps per pass (launch_multi) and the band height from the grid
  // and the kernel's residency (band_rows).  PH_HEAT_K, PH_HEAT_TY and
  // PH_HEAT_AHEAD override them.
  constexpr int64_t kcap = sizeof(T) == 4 ? 3 * V : 2 * V;
  int64_t kmax = kcap;
  int64_t ty = 0;
  int64_t ahead = 0;  // by steps per pass (launch_multi)
  if (const char *e = std::getenv("PH_HEAT_K")) kmax = std::min<int64_t>(kcap, std::max<int64_t>(1, std::atol(e)));
  if (const char *e = std::getenv("PH_HEAT_TY")) ty = std::max<int64_t>(8, std::atol(e));
  if (const char *f = std::getenv("PH_HEAT_AHEAD")) ahead = std::atol(f);
  HeatArgs a;
  a.ny = ny;
  a.nx = nx;
  a.nstrips = a.ty = a.nbands = a.nblocks = a.per_xcd = 0;  // set per pass by launch_ahead
  constexpr uintptr_t kAlign = sizeof(typename Vec<T>::type);
  const int vec_ok = (nx % V == 0) && ((uintptr_t)u % kAlign == 0) && ((uintptr_t)tmp % kAlign == 0);
  const T rr = (T)r;
  T *cur = u, *nxt = tmp;
  int64_t ei = 0;
  int rc;
  // The fused form needs one read-back of max|u| (a synchronisation), so
  // short runs keep the separate multiply and subtract.
  const bool try_fused = allow_fused && steps >= 32;
  bool fused = false;
  int64_t fused_until = 0;
  if (energy && every > 0) {
    rc = fp64_sum<T, T, 1>(cur, cur, ny * nx, energy + ei++, part_buf, part_cap);
    if (rc) return rc;
  }
  // Steps run kmax at a time up to each energy checkpoint; the remainder
  // before a checkpoint runs as one shorter pass.
  int64_t t = 0;
  while (t < steps) {
    int64_t until = steps;
    if (energy && every > 0) until = std::min<int64_t>(steps, (t / every + 1) * every);
    int k = (int)std::min<int64_t>(kmax, until - t);
    if (k > 8 && k < 12) k = 8;  // passes of more than 8 steps come only in twelves
    if (try_fused && t >= fused_until) {
      int64_t span = 0;
      rc = fused_safe<T>(cur, ny * nx, rr, &fused, &span);
      if (rc) return rc;
      fused_until = fused ? t + span : steps;
    }
    rc = ph::launch_gate("heat step");
    if (rc) return rc;
    rc = launch_k<T>(k, fused, cur, nxt, a, ty, rr, vec_ok, (int)ahead);
    if (rc) return rc;
    T *sw = cur;
    cur = nxt;
    nxt = sw;
    t += k;
    if (energy && every > 0 && t % every == 0) {
      rc = fp64_sum<T, T, 1>(cur, cur, ny * nx, energy + ei++, part_buf, part_cap);
      if (rc) return rc;
    }
  }
  if (cur != u) PH_HIP(hipMemcpyAsync(u, cur, sizeof(T) * (size_t)(ny * nx), hipMemcpyDeviceToDevice, ph::stream()));
  return PH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// The allocation-free heat run (ph_heat2d_noalloc) keeps its FP64 energy
// partials in this fixed device array instead of library scratch: 2048
// groups of one compensated pair, the most fp64_sum ever uses.  Its device
// address is looked up once per device when the library is initialised on
// it; host code uses that address, never the array's name.  Inside namespace
// ph the bare name would find the function ph::scratch (library scratch), so
// the lookup names the array as ::scratch.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[2048 * sizeof(D2)];

namespace {
constexpr int kMaxDevices = 64;
void *g_noalloc_scratch[kMaxDevices] = {nullptr};
}  // namespace

namespace ph {

int heat_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(PH_E_ARGUMENT, "device %d beyond the %d the fixed heat scratch tracks", device, kMaxDevices);
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(::scratch));
  if (e != hipSuccess || !p)
    return set_error(PH_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_noalloc_scratch[device] = p;
  return PH_OK;
}

}  // namespace ph

extern "C" {

int ph_heat2d_noalloc(int32_t dtype, void *u, void *tmp, int64_t ny, int64_t nx, double r, int64_t steps,
                      int64_t energy_every, double *energy_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (ny <= 0 || nx <= 0) return ph::set_error(PH_E_DIMENSION, "heat grid %lld x %lld", (long long)ny, (long long)nx);
  if (steps < 0) return ph::set_error(PH_E_ARGUMENT, "negative step count");
  if (!u || !tmp || u == tmp) return ph::set_error(PH_E_ARGUMENT, "heat2d needs two distinct buffers");
  if (energy_every > 0 && !energy_out) return ph::set_error(PH_E_ARGUMENT, "energy_out is null");
  const int dev = ph::device();
  void *sp = dev >= 0 && dev < kMaxDevices ? g_noalloc_scratch[dev] : nullptr;
  if (!sp) return ph::set_error(PH_E_NOINIT, "the fixed heat scratch of device %d was not set up", dev);
  const size_t cap = sizeof(::scratch);
  if (dtype == PH_F32)
    return heat_run<float>((float *)u, (float *)tmp, ny, nx, r, steps, energy_every, energy_out, sp, cap, false);
  if (dtype == PH_F64)
    return heat_run<double>((double *)u, (double *)tmp, ny, nx, r, steps, energy_every, energy_out, sp, cap, false);
  return ph::set_error(PH_E_ARGUMENT, "heat2d: dtype %d (F32 or F64)", dtype);
}

int ph_heat2d(int32_t dtype, void *u, void *tmp, int64_t ny, int64_t nx, double r, int64_t steps,
              int64_t energy_every, double *energy_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (ny <= 0 || nx <= 0) return ph::set_error(PH_E_DIMENSION, "heat grid %lld x %lld", (long long)ny, (long long)nx);
  if (steps < 0) return ph::set_error(PH_E_ARGUMENT, "negative step count");
  if (!u || !tmp || u == tmp) return ph::set_error(PH_E_ARGUMENT, "heat2d needs two distinct buffers");
  if (energy_every > 0 && !energy_out) return ph::set_error(PH_E_ARGUMENT, "energy_out is null");
  if (dtype == PH_F32) return heat_run<float>((float *)u, (float *)tmp, ny, nx, r, steps, energy_every, energy_out);
  if (dtype == PH_F64) return heat_run<double>((double *)u, (double *)tmp, ny, nx, r, steps, energy_every, energy_out);
  return ph::set_error(PH_E_ARGUMENT, "heat2d: dtype %d (F32 or F64)", dtype);
}

int ph_l2_diff(int32_t dtype_a, const void *a, int32_t dtype_b, const void *b, int64_t n, double *out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (n < 0 || !out) return ph::set_error(PH_E_ARGUMENT, "ph_l2_diff arguments");
  if (dtype_a == PH_F32 && dtype_b == PH_F64) return fp64_sum<float, double, 2>((const float *)a, (const double *)b, n, out);
  if (dtype_a == PH_F64 && dtype_b == PH_F64) return fp64_sum<double, double, 2>((const double *)a, (const double *)b, n, out);
  if (dtype_a == PH_F32 && dtype_b == PH_F32) return fp64_sum<float, float, 2>((const float *)a, (const float *)b, n, out);
  if (dtype_a == PH_F64 && dtype_b == PH_F32) return fp64_sum<double, float, 2>((const double *)a, (const float *)b, n, out);
  return ph::set_error(PH_E_ARGUMENT, "ph_l2_diff: dtypes %d, %d", dtype_a, dtype_b);
}

}  // extern "C"
