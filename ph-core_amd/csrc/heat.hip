// Explicit 2-D heat equation (5-point FTCS, Dirichlet zero outside the grid)
// and the FP64 reductions around it (total energy, relative L2 distance).
//
// Step kernel: a wavefront owns a strip of 64*VEC consecutive columns and
// sweeps TY rows top to bottom, keeping the north, centre and south rows of
// its strip in registers (each cell of the grid is read from memory once per
// step, plus one row above and below each TY-row band).  West/east neighbours
// come from the adjacent lane by a cross-lane shuffle; the two cells beyond
// the strip's edges are loaded by lanes 0 and 63.  Loads are VEC-wide
// (16 bytes per lane) and fully coalesced.  Workgroups are numbered so that
// strips adjacent in the grid run on the same XCD (blocks are dispatched to
// the 8 XCDs round-robin), which keeps the strip-edge loads in that XCD's L2.
//
// The cell update keeps the operation order of examples/heat_equation.cr:47
// and :50 extended to two axes, with no contraction into FMAs:
//   d = (((uN - 4u) + uS) + uW) + uE;  u' = u + d * r
#include <algorithm>
#include <cstdlib>
#include <cstring>

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

struct HeatArgs {
  int64_t ny, nx;
  int64_t ty;        // rows per band
  int64_t nstrips;   // strips per row band
  int64_t nbands;
  int64_t nblocks;   // workgroups
  int64_t per_xcd;   // workgroups per XCD slice
};

// Row y of the strip: VEC values per lane plus the west/east halo cells.
template <typename T>
__device__ __forceinline__ void load_row(const T *__restrict__ u, const HeatArgs &a, int64_t y,
                                         int64_t x0, int lane, bool vec_ok, T *v, T *west,
                                         T *east) {
  constexpr int V = Vec<T>::n;
  const int64_t x = x0 + (int64_t)lane * V;
  if (y < 0 || y >= a.ny) {
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = T(0);
    *west = T(0);
    *east = T(0);
    return;
  }
  const T *row = u + y * a.nx;
  if (vec_ok && x + V <= a.nx) {
    unpack<T>(*(const typename Vec<T>::type *)(row + x), v);
  } else {
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = (x + i < a.nx) ? row[x + i] : T(0);
  }
  *west = (lane == 0 && x0 > 0) ? row[x0 - 1] : T(0);
  const int64_t xe = x0 + 64 * V;
  *east = (lane == 63 && xe < a.nx) ? row[xe] : T(0);
}

template <typename T>
__global__ __launch_bounds__(64 * kWaves) void k_heat_step(const T *__restrict__ u,
                                                           T *__restrict__ out, HeatArgs a, T r,
                                                           int vec_ok) {
  constexpr int V = Vec<T>::n;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // XCD-aware renumbering: block b runs on XCD b % 8; give each XCD a
  // contiguous range of logical blocks.
  const int64_t b = blockIdx.x;
  const int64_t lb = (b % 8) * a.per_xcd + b / 8;
  if (lb >= a.nblocks) return;
  const int64_t strip = (lb % ((a.nstrips + kWaves - 1) / kWaves)) * kWaves + wave;
  const int64_t band = lb / ((a.nstrips + kWaves - 1) / kWaves);
  if (strip >= a.nstrips) return;
  const int64_t x0 = strip * 64 * V;
  const int64_t y0 = band * a.ty;
  const int64_t y1 = (y0 + a.ty < a.ny) ? y0 + a.ty : a.ny;
  T n[V], c[V], cw, ce, dummy_w, dummy_e;
  load_row<T>(u, a, y0 - 1, x0, lane, vec_ok, n, &dummy_w, &dummy_e);
  load_row<T>(u, a, y0, x0, lane, vec_ok, c, &cw, &ce);
  const T four = T(4);
  const int64_t x = x0 + (int64_t)lane * V;
  // kAhead south rows are loaded together before any of them is used, so a
  // wavefront keeps kAhead row loads in flight instead of one.
  constexpr int kAhead = 4;
  for (int64_t y = y0; y < y1; y += kAhead) {
    T s[kAhead][V], sw[kAhead], se[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      if (y + 1 + k <= y1) load_row<T>(u, a, y + 1 + k, x0, lane, vec_ok, s[k], &sw[k], &se[k]);
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      if (y + k >= y1) break;
      // west neighbour of c[0] is the previous lane's c[V-1]; east of c[V-1]
      // is the next lane's c[0]
      T wl = __shfl_up(c[V - 1], 1, 64);
      T er = __shfl_down(c[0], 1, 64);
      if (lane == 0) wl = cw;
      if (lane == 63) er = ce;
      T o[V];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const T w = i == 0 ? wl : c[i - 1];
        const T e = i == V - 1 ? er : c[i + 1];
        const T d = (((n[i] - four * c[i]) + s[k][i]) + w) + e;
        o[i] = c[i] + d * r;
      }
      T *orow = out + (y + k) * a.nx;
      if (vec_ok && x + V <= a.nx) {
        *(typename Vec<T>::type *)(orow + x) = pack<T>(o);
      } else {
#pragma unroll
        for (int i = 0; i < V; ++i)
          if (x + i < a.nx) orow[x + i] = o[i];
      }
#pragma unroll
      for (int i = 0; i < V; ++i) {
        n[i] = c[i];
        c[i] = s[k][i];
      }
      cw = sw[k];
      ce = se[k];
    }
  }
}

// ---------------------------------------------------------------------------
// Two time steps per pass (temporal blocking).  The wavefront sweeps its strip
// keeping three rows of level t0 (read from memory) and three rows of level
// t1 (computed) in registers; level t2 is written.  Level-t1 values are also
// needed one column beyond each strip edge, so lanes 0 and 63 carry two halo
// columns of t0 and one of t1.  Every t1 and t2 value is computed with the
// single-step formula, so the result is bit-identical to two single steps;
// cells outside the grid stay 0 at every level.
// ---------------------------------------------------------------------------
template <typename T>
struct Row0 {  // a t0 row of the strip plus two halo columns on each side
  T v[Vec<T>::n];
  T w1, w2, e1, e2;  // x0-1, x0-2, x0+W, x0+W+1
};
template <typename T>
struct Row1 {  // a t1 row plus one halo column on each side
  T v[Vec<T>::n];
  T w, e;  // x0-1, x0+W
};

template <typename T>
__device__ __forceinline__ void load_row0(const T *__restrict__ u, const HeatArgs &a, int64_t y, int64_t x0,
                                          int lane, bool vec_ok, Row0<T> &r) {
  constexpr int V = Vec<T>::n;
  const int64_t x = x0 + (int64_t)lane * V;
  r.w1 = r.w2 = r.e1 = r.e2 = T(0);
  if (y < 0 || y >= a.ny) {
#pragma unroll
    for (int i = 0; i < V; ++i) r.v[i] = T(0);
    return;
  }
  const T *row = u + y * a.nx;
  if (vec_ok && x + V <= a.nx) {
    unpack<T>(*(const typename Vec<T>::type *)(row + x), r.v);
  } else {
#pragma unroll
    for (int i = 0; i < V; ++i) r.v[i] = (x + i < a.nx) ? row[x + i] : T(0);
  }
  const int64_t xe = x0 + 64 * V;
  if (lane == 0) {
    r.w1 = x0 - 1 >= 0 ? row[x0 - 1] : T(0);
    r.w2 = x0 - 2 >= 0 ? row[x0 - 2] : T(0);
  }
  if (lane == 63) {
    r.e1 = xe < a.nx ? row[xe] : T(0);
    r.e2 = xe + 1 < a.nx ? row[xe + 1] : T(0);
  }
}

// t1 row from t0 rows n (above), c, s (below); row index y1row.
template <typename T>
__device__ __forceinline__ void step_row(const Row0<T> &n, const Row0<T> &c, const Row0<T> &s, T r,
                                         const HeatArgs &a, int64_t yrow, int64_t x0, int lane, Row1<T> &o) {
  constexpr int V = Vec<T>::n;
  const T four = T(4);
  if (yrow < 0 || yrow >= a.ny) {
#pragma unroll
    for (int i = 0; i < V; ++i) o.v[i] = T(0);
    o.w = o.e = T(0);
    return;
  }
  T wl = __shfl_up(c.v[V - 1], 1, 64);
  T er = __shfl_down(c.v[0], 1, 64);
  if (lane == 0) wl = c.w1;
  if (lane == 63) er = c.e1;
  const int64_t x = x0 + (int64_t)lane * V;
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const T w = i == 0 ? wl : c.v[i - 1];
    const T e = i == V - 1 ? er : c.v[i + 1];
    const T d = (((n.v[i] - four * c.v[i]) + s.v[i]) + w) + e;
    o.v[i] = (x + i < a.nx) ? c.v[i] + d * r : T(0);
  }
  o.w = T(0);
  o.e = T(0);
  if (lane == 0 && x0 - 1 >= 0) {  // column x0-1: west x0-2, east x0 (own v[0])
    const T d = (((n.w1 - four * c.w1) + s.w1) + c.w2) + c.v[0];
    o.w = c.w1 + d * r;
  }
  const int64_t xe = x0 + 64 * V;
  if (lane == 63 && xe < a.nx) {  // column x0+W: west x0+W-1 (own v[V-1]), east x0+W+1
    const T d = (((n.e1 - four * c.e1) + s.e1) + c.v[V - 1]) + c.e2;
    o.e = c.e1 + d * r;
  }
}

template <typename T, int kAhead>
__global__ __launch_bounds__(64 * kWaves) void k_heat_step2(const T *__restrict__ u,
                                                            T *__restrict__ out, HeatArgs a, T r,
                                                            int vec_ok) {
  constexpr int V = Vec<T>::n;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t b = blockIdx.x;
  const int64_t lb = (b % 8) * a.per_xcd + b / 8;
  if (lb >= a.nblocks) return;
  const int64_t gpb = (a.nstrips + kWaves - 1) / kWaves;
  const int64_t strip = (lb % gpb) * kWaves + wave;
  const int64_t band = lb / gpb;
  if (strip >= a.nstrips) return;
  const int64_t x0 = strip * 64 * V;
  const int64_t y0 = band * a.ty;
  const int64_t y1 = (y0 + a.ty < a.ny) ? y0 + a.ty : a.ny;
  const int64_t x = x0 + (int64_t)lane * V;
  const T four = T(4);
  // start-up: t0 rows y0-2 .. y0+1, t1 rows y0-1, y0
  Row0<T> t0a, t0b, t0c, t0d;
  load_row0<T>(u, a, y0 - 2, x0, lane, vec_ok, t0a);
  load_row0<T>(u, a, y0 - 1, x0, lane, vec_ok, t0b);
  load_row0<T>(u, a, y0, x0, lane, vec_ok, t0c);
  load_row0<T>(u, a, y0 + 1, x0, lane, vec_ok, t0d);
  Row1<T> P, Q;
  step_row<T>(t0a, t0b, t0c, r, a, y0 - 1, x0, lane, P);
  step_row<T>(t0b, t0c, t0d, r, a, y0, x0, lane, Q);
  Row0<T> A = t0c, B = t0d;
  for (int64_t y = y0; y < y1; y += kAhead) {
    Row0<T> C[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
      if (y + k < y1) load_row0<T>(u, a, y + 2 + k, x0, lane, vec_ok, C[k]);
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      if (y + k >= y1) break;
      Row1<T> R;
      step_row<T>(A, B, C[k], r, a, y + k + 1, x0, lane, R);
      // level t2 row y+k from t1 rows P (above), Q, R (below)
      T wl = __shfl_up(Q.v[V - 1], 1, 64);
      T er = __shfl_down(Q.v[0], 1, 64);
      if (lane == 0) wl = Q.w;
      if (lane == 63) er = Q.e;
      T o[V];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const T w = i == 0 ? wl : Q.v[i - 1];
        const T e = i == V - 1 ? er : Q.v[i + 1];
        const T d = (((P.v[i] - four * Q.v[i]) + R.v[i]) + w) + e;
        o[i] = Q.v[i] + d * r;
      }
      T *orow = out + (y + k) * a.nx;
      if (vec_ok && x + V <= a.nx) {
        *(typename Vec<T>::type *)(orow + x) = pack<T>(o);
      } else {
#pragma unroll
        for (int i = 0; i < V; ++i)
          if (x + i < a.nx) orow[x + i] = o[i];
      }
      A = B;
      B = C[k];
      P = Q;
      Q = R;
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

template <typename TA, typename TB, int K>
int fp64_sum(const TA *a, const TB *b, int64_t n, double *out) {
  const int64_t nparts_max = 2048;
  int64_t chunk = (n + nparts_max - 1) / nparts_max;
  if (chunk < 4096) chunk = 4096;
  const int64_t nparts = (n + chunk - 1) / chunk;
  void *part;
  int rc = ph::scratch((size_t)(nparts > 0 ? nparts : 1) * K * sizeof(D2), &part, 2);
  if (rc) return rc;
  rc = ph::launch_gate("fp64 sum");
  if (rc) return rc;
  if (nparts > 0)
    hipLaunchKernelGGL((k_sum_pass1<TA, TB, K>), dim3((unsigned)nparts), dim3(256), 0, ph::stream(),
                       a, b, n, chunk, (D2 *)part);
  hipLaunchKernelGGL((k_sum_pass2<K>), dim3(1), dim3(256), 0, ph::stream(), (const D2 *)part,
                     nparts, out);
  return ph::check_launch("k_sum");
}

template <typename T>
int heat_run(T *u, T *tmp, int64_t ny, int64_t nx, double r, int64_t steps, int64_t every,
             double *energy) {
  constexpr int V = Vec<T>::n;
  HeatArgs a;
  a.ny = ny;
  a.nx = nx;
  a.nstrips = (nx + 64 * V - 1) / (64 * V);
  // Band height and row prefetch depth of the two-step kernel, per precision
  // (measured on MI355X at 8192^2: FP32 32 rows / no prefetch, FP64 16 rows /
  // 2 rows ahead); PH_HEAT_TY and PH_HEAT_AHEAD override them for tuning.
  int64_t ty_env = sizeof(T) == 4 ? 32 : 16;
  int64_t ahead_env = sizeof(T) == 4 ? 1 : 2;
  if (const char *e = std::getenv("PH_HEAT_TY")) ty_env = std::max<int64_t>(8, std::atol(e));
  if (const char *f = std::getenv("PH_HEAT_AHEAD")) ahead_env = std::atol(f);
  a.ty = ty_env;
  a.nbands = (ny + a.ty - 1) / a.ty;
  const int64_t groups_per_band = (a.nstrips + kWaves - 1) / kWaves;
  a.nblocks = groups_per_band * a.nbands;
  a.per_xcd = (a.nblocks + 7) / 8;
  const int64_t grid = a.per_xcd * 8;
  constexpr uintptr_t kAlign = sizeof(typename Vec<T>::type);
  const int vec_ok = (nx % V == 0) && ((uintptr_t)u % kAlign == 0) && ((uintptr_t)tmp % kAlign == 0);
  const T rr = (T)r;
  T *cur = u, *nxt = tmp;
  int64_t ei = 0;
  int rc;
  if (energy && every > 0) {
    rc = fp64_sum<T, T, 1>(cur, cur, ny * nx, energy + ei++);
    if (rc) return rc;
  }
  // Steps run two at a time (k_heat_step2) up to each energy checkpoint; an
  // odd remainder runs as a single step.
  int64_t t = 0;
  while (t < steps) {
    int64_t until = steps;
    if (energy && every > 0) until = std::min<int64_t>(steps, (t / every + 1) * every);
    const bool pair = until - t >= 2;
    rc = ph::launch_gate("heat step");
    if (rc) return rc;
    if (pair) {
      if (ahead_env == 4)
        hipLaunchKernelGGL((k_heat_step2<T, 4>), dim3((unsigned)grid), dim3(64 * kWaves), 0, ph::stream(),
                           (const T *)cur, nxt, a, rr, vec_ok);
      else if (ahead_env == 1)
        hipLaunchKernelGGL((k_heat_step2<T, 1>), dim3((unsigned)grid), dim3(64 * kWaves), 0, ph::stream(),
                           (const T *)cur, nxt, a, rr, vec_ok);
      else
        hipLaunchKernelGGL((k_heat_step2<T, 2>), dim3((unsigned)grid), dim3(64 * kWaves), 0, ph::stream(),
                           (const T *)cur, nxt, a, rr, vec_ok);
      rc = ph::check_launch("k_heat_step2");
    } else {
      hipLaunchKernelGGL(k_heat_step<T>, dim3((unsigned)grid), dim3(64 * kWaves), 0, ph::stream(),
                         (const T *)cur, nxt, a, rr, vec_ok);
      rc = ph::check_launch("k_heat_step");
    }
    if (rc) return rc;
    T *sw = cur;
    cur = nxt;
    nxt = sw;
    t += pair ? 2 : 1;
    if (energy && every > 0 && t % every == 0) {
      rc = fp64_sum<T, T, 1>(cur, cur, ny * nx, energy + ei++);
      if (rc) return rc;
    }
  }
  if (cur != u) PH_HIP(hipMemcpyAsync(u, cur, sizeof(T) * (size_t)(ny * nx), hipMemcpyDeviceToDevice, ph::stream()));
  return PH_OK;
}

}  // namespace

extern "C" {

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
