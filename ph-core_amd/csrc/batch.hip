// Batched small arrays: for every array k of a batch (Float32, rank 1..3,
// packed row-major into one buffer),
//     y_k = x_k + s_k                  (NArray#map / scalar operator)
//     sum_k = x_k.sum                  (Enumerable#sum)
//     z_k = x_k.view(0..2.., ...).to_narr   (step-2 slice on every axis)
// in one pass over x: a persistent kernel whose wavefronts run work items
// (an array, or a chunk of a large one) from a device-side table.  A
// wavefront streams its item's elements once (16-byte loads of x, 16-byte
// non-temporal stores of y, FP32 Neumaier sums), 2048 elements at a time,
// keeps each step's values in LDS, gathers from there the slice elements
// sourced from the step and writes them to z contiguously.  Per-item partial sums land in fixed slots and a second
// kernel folds each array's slots in item order, so the sums do not depend
// on which wavefront ran which item.  The kernels are captured once in a HIP
// graph and replayed.
//
// Schedule: items are chunks of at most `chunk` elements, ordered largest
// first; wavefronts take them either round-robin (static) or from an atomic
// queue (dynamic).  The autotuner times (chunk, workgroups per CU, static or
// dynamic) on the batch itself and keeps the fastest, cached by a histogram
// of array sizes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "common.hpp"

namespace {

struct BArray {  // host-side description of one array
  int64_t xoff, zoff;   // element offsets into the packed x/y and z buffers
  int64_t n;            // elements
  int32_t rank, pad_;
  uint32_t shp[3];      // shape (rank padded on the left with 1)
  uint32_t zshp[3];     // slice shape: ceil(shp/2)
  float scalar, pad2_;
};

// One work item, in dispatch order: everything its wavefront needs.
struct BRec {
  int64_t xoff, zoff;               // the array's offsets in x/y and z
  uint32_t lo, hi;                  // element range [lo, hi) of the array
  uint32_t zlo, zhi;                // slice elements whose source lies in [lo, hi)
  uint32_t shp[3], zshp[3];
  uint32_t m_c, l_c, m_b, l_b;      // magic divisors for shp[2], shp[1]
  uint32_t mz_c, lz_c, mz_b, lz_b;  // magic divisors for zshp[2], zshp[1]
  float scalar;
  int32_t slot;                     // partial-sum slot (array-major item index)
};
static_assert(sizeof(BRec) == 96, "BRec layout");

struct Batch {
  int n = 0;
  std::vector<BArray> arrays;
  int64_t total = 0, ztotal = 0;
  // device buffers owned by the batch
  BRec *d_recs = nullptr;         // work items in dispatch order
  int32_t *d_first = nullptr;     // first slot of each array (n+1)
  float2 *d_partial = nullptr;
  int32_t *d_counter = nullptr;
  int64_t nitems = 0;
  int64_t items_cap = 0;
  // schedule
  int64_t chunk = 0;
  int grid = 0;
  int dynamic = 1;
  int sub = 2048;  // elements per streaming step of a wavefront (1024 or 2048)
  // caller buffers
  const float *x = nullptr;
  float *y = nullptr, *z = nullptr;
  double *sums = nullptr;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void magic_u32(uint32_t d, uint32_t *m, uint32_t *l) {
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  *l = s;
  *m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t m, uint32_t l) {
  return (uint32_t)(((uint64_t)__umulhi(n, m) + n) >> l);
}

__device__ __forceinline__ void neu_add(float &s, float &c, float v) {
  const float t = s + v;
  const float d = (fabsf(s) >= fabsf(v)) ? ((s - t) + v) : ((v - t) + s);
  c += d;
  s = t;
}

__device__ __forceinline__ void two_sum_fold(float &s, float &c, float s2, float c2) {
  const float t = s + s2;
  const float bb = t - s;
  const float err = (s - (t - bb)) + (s2 - bb);
  s = t;
  c = (c + c2) + err;
}

// Butterfly fold of the 64 lanes' (sum, compensation) pairs with TwoSum.
__device__ __forceinline__ void wave_fold(float &s, float &c) {
  for (int off = 32; off >= 1; off >>= 1) {
    const float s2 = __shfl_xor(s, off, 64);
    const float c2 = __shfl_xor(c, off, 64);
    two_sum_fold(s, c, s2, c2);
  }
}

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void neu_add4(float &s, float &c, f4 v) {
  neu_add(s, c, v.x);
  neu_add(s, c, v.y);
  neu_add(s, c, v.z);
  neu_add(s, c, v.w);
}

// Slice elements whose source index is < e (e < the array's size): the
// sources (2a, 2b, 2c) increase with the slice index, so the slice elements
// sourced from a range of x form a contiguous range of z.
__device__ __forceinline__ uint32_t slice_before(const BRec &R, uint32_t e) {
  const uint32_t t = fdiv(e, R.m_c, R.l_c);
  const uint32_t k = e - t * R.shp[2];
  const uint32_t i = fdiv(t, R.m_b, R.l_b);
  const uint32_t j = t - i * R.shp[1];
  uint32_t cnt = ((i + 1) >> 1) * R.zshp[1] * R.zshp[2];  // slice planes of even i' < i
  if ((i & 1u) == 0) {
    cnt += ((j + 1) >> 1) * R.zshp[2];                   // slice rows of even j' < j
    if ((j & 1u) == 0) cnt += (k + 1) >> 1;               // even k' < k
  }
  return cnt;
}

// Persistent: each wavefront runs work items until none are left.  dynamic:
// items come from a queue (one atomic per item, lane 0); static: wavefront w
// of W takes dispatch positions w, w + W, w + 2W, ... of the largest-first
// order, which needs no atomics.  An item is processed kSub elements at a
// time (a template parameter, 1024 or 2048): lane l takes elements
// base + 4l + 256m (m = 0 .. kSub/256 - 1) as 16-byte vectors (items start at
// multiples of 4 and every array is padded to 4 elements in the packed
// layout, so a vector never leaves x; a partial last vector is stored
// element by element) and also stores them to the wavefront's LDS copy of
// the step; the wavefront then gathers the slice elements sourced from
// [base, base + kSub) out of that copy and writes them to z, so x is read
// from memory once.
template <uint32_t kSub>
__global__ __launch_bounds__(256) void k_batch(const BRec *__restrict__ recs, int32_t nitems, int32_t *counter,
                                               int32_t dynamic, const float *__restrict__ x,
                                               float *__restrict__ y, float *__restrict__ z,
                                               float2 *__restrict__ partial) {
  const int lane = threadIdx.x & 63;
  const int32_t nwaves = (int32_t)gridDim.x * 4;
  int32_t q = (int32_t)blockIdx.x * 4 + (int32_t)(threadIdx.x >> 6);
  // each wavefront's copy of its current step of x, for the slice gather
  __shared__ f4 sx_all[4][kSub / 4];
  f4 *sx = sx_all[threadIdx.x >> 6];
  if (dynamic) {
    int32_t t = 0;
    if (lane == 0) t = atomicAdd(counter, 1);
    q = __builtin_amdgcn_readfirstlane(t);
  }
  while (q < nitems) {  // every wavefront leaves once the items are exhausted
    const BRec R = recs[q];
    const float *__restrict__ xa = x + R.xoff;
    float *__restrict__ ya = y + R.xoff;
    float *__restrict__ za = z + R.zoff;
    float s = 0.f, c = 0.f;
    uint32_t zlo = R.zlo;
    for (uint32_t base = R.lo; base < R.hi; base += kSub) {
      const uint32_t end = min(R.hi, base + kSub);
      constexpr int kVecs = (int)(kSub / 256);
      f4 v[kVecs];
#pragma unroll
      for (int m = 0; m < kVecs; ++m) {
        const uint32_t e = base + 4u * (uint32_t)lane + 256u * m;
        if (e < end) v[m] = *(const f4 *)(xa + e);
      }
      // (positions past `end` hold stale values; the gather never reads them)
#pragma unroll
      for (int m = 0; m < kVecs; ++m) sx[lane + 64 * m] = v[m];
#pragma unroll
      for (int m = 0; m < kVecs; ++m) {
        const uint32_t e = base + 4u * (uint32_t)lane + 256u * m;
        const f4 r = v[m] + R.scalar;
        if (e + 4 <= end) {
          __builtin_nontemporal_store(r, (f4 *)(ya + e));
          neu_add4(s, c, v[m]);
        } else if (e < end) {
          const uint32_t cnt = end - e;  // 1..3
          ya[e] = r.x;
          neu_add(s, c, v[m].x);
          if (cnt > 1) {
            ya[e + 1] = r.y;
            neu_add(s, c, v[m].y);
          }
          if (cnt > 2) {
            ya[e + 2] = r.z;
            neu_add(s, c, v[m].z);
          }
        }
      }
      // slice elements in z order: source (2a, 2b, 2c) of z (a, b, c), which
      // lies in [base, end): read from the wavefront's LDS copy of the step;
      // four per lane at a time, their reads issued before the stores
      __builtin_amdgcn_wave_barrier();
      const float *sxf = (const float *)sx;
      const uint32_t zhi = end == R.hi ? R.zhi : slice_before(R, end);
      for (uint32_t z0 = zlo; z0 < zhi; z0 += 256) {
        float g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t zi = z0 + 64u * u + (uint32_t)lane;
          if (zi < zhi) {
            const uint32_t t = fdiv(zi, R.mz_c, R.lz_c);
            const uint32_t cc = zi - t * R.zshp[2];
            const uint32_t aa = fdiv(t, R.mz_b, R.lz_b);
            const uint32_t bb = t - aa * R.zshp[1];
            g[u] = sxf[((2 * aa) * R.shp[1] + 2 * bb) * R.shp[2] + 2 * cc - base];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t zi = z0 + 64u * u + (uint32_t)lane;
          if (zi < zhi) za[zi] = g[u];
        }
      }
      zlo = zhi;
      __builtin_amdgcn_wave_barrier();  // the next step overwrites the LDS copy
    }
    wave_fold(s, c);
    if (lane == 0) partial[R.slot] = make_float2(s, c);
    if (dynamic) {
      int32_t t = 0;
      if (lane == 0) t = atomicAdd(counter, 1);
      q = __builtin_amdgcn_readfirstlane(t);
    } else {
      q += nwaves;
    }
  }
}

// Per array (one wavefront each), fold its items' partials with a fixed
// tree: lane l folds a run of consecutive slots in order, then the lanes are
// combined by the butterfly; the result depends only on the slot values.
__global__ __launch_bounds__(256) void k_batch_final(const int32_t *__restrict__ first, int n,
                                                     const float2 *__restrict__ partial, double *sums) {
  const int k = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= n) return;  // uniform per wavefront
  const int32_t b = first[k], e = first[k + 1];
  const int32_t per = (e - b + 63) / 64;
  const int32_t lo = b + lane * per, hi = min(e, lo + per);
  float s = 0.f, c = 0.f;
  for (int32_t i = lo; i < hi; ++i) two_sum_fold(s, c, partial[i].x, partial[i].y);
  wave_fold(s, c);
  if (lane == 0) sums[k] = (double)s + (double)c;
}

__global__ void k_reset(int32_t *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

// --- per-array launches (the unbatched comparison) --------------------------
__global__ __launch_bounds__(256) void k_one_add(const float *__restrict__ x, float *__restrict__ y, int64_t off,
                                                 int64_t n, float s) {
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) y[off + e] = x[off + e] + s;
}
__global__ __launch_bounds__(256) void k_one_sum(const float *__restrict__ x, int64_t off, int64_t n, double *out) {
  __shared__ float ss[4], cc[4];
  float s = 0.f, c = 0.f;
  for (int64_t e = threadIdx.x; e < n; e += 256) neu_add(s, c, x[off + e]);
  wave_fold(s, c);
  if ((threadIdx.x & 63) == 0) {
    ss[threadIdx.x >> 6] = s;
    cc[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = ss[0], b = cc[0];
    for (int w = 1; w < 4; ++w) two_sum_fold(a, b, ss[w], cc[w]);
    *out = (double)a + (double)b;
  }
}
__global__ __launch_bounds__(256) void k_one_slice(const float *__restrict__ x, float *__restrict__ z, BArray A,
                                                   int64_t nz) {
  for (int64_t zi = blockIdx.x * 256 + threadIdx.x; zi < nz; zi += (int64_t)gridDim.x * 256) {
    const uint32_t k = (uint32_t)(zi % A.zshp[2]);
    const uint32_t t = (uint32_t)(zi / A.zshp[2]);
    const uint32_t j = t % A.zshp[1];
    const uint32_t i = t / A.zshp[1];
    const int64_t src = ((int64_t)(2 * i) * A.shp[1] + 2 * j) * A.shp[2] + 2 * k;
    z[A.zoff + zi] = x[A.xoff + src];
  }
}

// ---------------------------------------------------------------------------
// Slice elements of an array of shape shp whose source index is < e:
// sources are row-major indices of (2a, 2b, 2c), increasing with the slice
// index, so the slice elements sourced from [lo, hi) are a contiguous range.
int64_t slice_before(const BArray &A, int64_t e) {
  const int64_t S2 = A.shp[2], S1 = A.shp[1];
  const int64_t Z2 = A.zshp[2], Z1 = A.zshp[1], Z0 = A.zshp[0];
  if (e >= A.n) return Z0 * Z1 * Z2;
  const int64_t k = e % S2, t = e / S2, j = t % S1, i = t / S1;
  int64_t cnt = ((i + 1) / 2) * Z1 * Z2;  // whole slice planes of even i' < i
  if (i % 2 == 0) {
    cnt += ((j + 1) / 2) * Z2;                // whole slice rows of even j' < j
    if (j % 2 == 0) cnt += (k + 1) / 2;       // even k' < k
  }
  return cnt;
}

int build_items(Batch *B, int64_t chunk) {
  std::vector<BRec> items;
  std::vector<int32_t> first((size_t)B->n + 1);
  for (int k = 0; k < B->n; ++k) {
    first[(size_t)k] = (int32_t)items.size();
    const BArray &A = B->arrays[(size_t)k];
    BRec r;
    std::memset(&r, 0, sizeof(r));
    r.xoff = A.xoff;
    r.zoff = A.zoff;
    for (int i = 0; i < 3; ++i) {
      r.shp[i] = A.shp[i];
      r.zshp[i] = A.zshp[i];
    }
    magic_u32(A.shp[2], &r.m_c, &r.l_c);
    magic_u32(A.shp[1], &r.m_b, &r.l_b);
    magic_u32(A.zshp[2], &r.mz_c, &r.lz_c);
    magic_u32(A.zshp[1], &r.mz_b, &r.lz_b);
    r.scalar = A.scalar;
    for (int64_t lo = 0; lo < A.n; lo += chunk) {
      r.lo = (uint32_t)lo;
      r.hi = (uint32_t)std::min(A.n, lo + chunk);
      r.zlo = (uint32_t)slice_before(A, lo);
      r.zhi = (uint32_t)slice_before(A, r.hi);
      r.slot = (int32_t)items.size();
      items.push_back(r);
    }
  }
  first[(size_t)B->n] = (int32_t)items.size();
  const int64_t ni = (int64_t)items.size();
  if (ni >= INT32_MAX) return ph::set_error(PH_E_ARGUMENT, "batch: too many work items (%lld)", (long long)ni);
  // dispatch order: largest first (ties in array order)
  std::stable_sort(items.begin(), items.end(),
                   [](const BRec &a, const BRec &b) { return a.hi - a.lo > b.hi - b.lo; });
  if (ni > B->items_cap) {
    if (B->d_recs) {
      PH_HIP(hipStreamSynchronize(ph::stream()));
      PH_HIP(hipFree(B->d_recs));
      PH_HIP(hipFree(B->d_partial));
      B->d_recs = nullptr;
      B->d_partial = nullptr;
      B->items_cap = 0;
    }
    PH_HIP(hipMalloc(&B->d_recs, sizeof(BRec) * (size_t)ni));
    PH_HIP(hipMalloc(&B->d_partial, sizeof(float2) * (size_t)ni));
    B->items_cap = ni;
  }
  PH_HIP(hipStreamSynchronize(ph::stream()));
  if (ni) PH_HIP(hipMemcpy(B->d_recs, items.data(), sizeof(BRec) * (size_t)ni, hipMemcpyHostToDevice));
  PH_HIP(hipMemcpy(B->d_first, first.data(), sizeof(int32_t) * first.size(), hipMemcpyHostToDevice));
  PH_HIP(hipMemset(B->d_counter, 0, 64));
  B->nitems = ni;
  B->chunk = chunk;
  return PH_OK;
}

int launch_batch(Batch *B) {
  if (B->n == 0) return PH_OK;  // nothing to launch
  hipStream_t st = ph::stream();
  if (B->dynamic) hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, B->d_counter);
  int sub = B->sub;
  if (const char *e = std::getenv("PH_BATCH_SUB")) sub = std::atoi(e);  // tuning override
  if (sub == 2048)
    hipLaunchKernelGGL(k_batch<2048>, dim3((unsigned)B->grid), dim3(256), 0, st, (const BRec *)B->d_recs,
                       (int32_t)B->nitems, B->d_counter, (int32_t)B->dynamic, B->x, B->y, B->z, B->d_partial);
  else
    hipLaunchKernelGGL(k_batch<1024>, dim3((unsigned)B->grid), dim3(256), 0, st, (const BRec *)B->d_recs,
                       (int32_t)B->nitems, B->d_counter, (int32_t)B->dynamic, B->x, B->y, B->z, B->d_partial);
  hipLaunchKernelGGL(k_batch_final, dim3((unsigned)((B->n + 3) / 4)), dim3(256), 0, st,
                     (const int32_t *)B->d_first, B->n, (const float2 *)B->d_partial, B->sums);
  return ph::check_launch("k_batch");
}

int capture(Batch *B) {
  if (B->exec) {
    PH_HIP(hipGraphExecDestroy(B->exec));
    B->exec = nullptr;
  }
  if (B->graph) {
    PH_HIP(hipGraphDestroy(B->graph));
    B->graph = nullptr;
  }
  PH_HIP(hipStreamSynchronize(ph::stream()));
  PH_HIP(hipStreamBeginCapture(ph::stream(), hipStreamCaptureModeThreadLocal));
  int rc = launch_batch(B);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ph::stream(), &g);
  if (rc) return rc;
  if (e != hipSuccess) return ph::set_error(PH_E_HIP, "graph capture: %s", hipGetErrorString(e));
  B->graph = g;
  PH_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
  return PH_OK;
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

struct Schedule {
  int64_t chunk;
  int grid;
  int dynamic;
};

std::map<std::string, Schedule> &tune_cache() {
  static std::map<std::string, Schedule> c;
  return c;
}

std::string histogram_key(const Batch *B) {
  int64_t h[40] = {0};
  for (const BArray &a : B->arrays) {
    int b = 0;
    while (b < 39 && (int64_t(1) << b) < a.n) ++b;
    ++h[b];
  }
  std::string k((const char *)h, sizeof(h));
  k.append((const char *)&B->n, sizeof(B->n));
  return k;
}

}  // namespace

extern "C" {

int ph_batch_destroy(void *handle);

int ph_batch_create(void **handle, int32_t n, const int32_t *ranks, const int64_t *shapes, const float *x,
                    const float *scalars, float *y, float *z, double *sums) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (!handle || n < 0 || (n > 0 && (!ranks || !shapes || !scalars)))
    return ph::set_error(PH_E_ARGUMENT, "ph_batch_create arguments");
  if (((uintptr_t)x | (uintptr_t)y) % 16 != 0)
    return ph::set_error(PH_E_ARGUMENT, "batch: x and y must be 16-byte aligned");
  Batch *B = new Batch();
  B->n = n;
  B->x = x;
  B->y = y;
  B->z = z;
  B->sums = sums;
  int64_t off = 0, zoff = 0;
  for (int k = 0; k < n; ++k) {
    BArray a;
    std::memset(&a, 0, sizeof(a));
    const int r = ranks[k];
    if (r < 1 || r > 3) {
      delete B;
      return ph::set_error(PH_E_DIMENSION, "batch array %d: rank %d (1..3)", k, r);
    }
    a.rank = r;
    int64_t s3[3] = {1, 1, 1};
    for (int i = 0; i < r; ++i) {
      const int64_t d = shapes[3 * k + i];
      if (d < 0 || d > (int64_t(1) << 30)) {
        delete B;
        return ph::set_error(PH_E_DIMENSION, "batch array %d: extent %lld", k, (long long)d);
      }
      if (d == 0) {  // the step-2 view of an empty axis raises in the reference
        delete B;
        return ph::set_error(PH_E_INDEX,
                             "Invalid index: At least one endpoint is negative after canonicalization. "
                             "(batch array %d, axis %d has extent 0)",
                             k, i);
      }
      s3[3 - r + i] = d;
    }
    a.n = s3[0] * s3[1] * s3[2];
    if (a.n >= (int64_t(1) << 31)) {
      delete B;
      return ph::set_error(PH_E_DIMENSION, "batch array %d: more than 2^31 elements", k);
    }
    for (int i = 0; i < 3; ++i) {
      a.shp[i] = (uint32_t)s3[i];
      a.zshp[i] = (uint32_t)((s3[i] + 1) / 2);
    }
    a.xoff = off;
    a.zoff = zoff;
    a.scalar = scalars[k];
    off += (a.n + 3) / 4 * 4;  // arrays start on 16-byte boundaries
    zoff += (int64_t)a.zshp[0] * a.zshp[1] * a.zshp[2];
    B->arrays.push_back(a);
  }
  B->total = off;
  B->ztotal = zoff;
  if (hipMalloc(&B->d_first, sizeof(int32_t) * (size_t)(n + 1)) != hipSuccess ||
      hipMalloc(&B->d_counter, 64) != hipSuccess) {
    ph_batch_destroy(B);
    return ph::set_error(PH_E_OOM, "batch tables");
  }
  B->grid = cus() * 4;
  int rc = build_items(B, 16384);
  if (rc) {
    ph_batch_destroy(B);
    return rc;
  }
  *handle = B;
  return PH_OK;
}

// Element counts the caller must allocate: packed x/y (aligned) and z.
int ph_batch_sizes(void *handle, int64_t *xy_elements, int64_t *z_elements, int64_t *offsets_x,
                   int64_t *offsets_z) {
  PH_GUARD();
  Batch *B = (Batch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  *xy_elements = B->total;
  *z_elements = B->ztotal;
  for (int k = 0; k < B->n; ++k) {
    if (offsets_x) offsets_x[k] = B->arrays[(size_t)k].xoff;
    if (offsets_z) offsets_z[k] = B->arrays[(size_t)k].zoff;
  }
  return PH_OK;
}

// Points the batch at (new) caller buffers; invalidates a captured graph.
int ph_batch_bind(void *handle, const float *x, float *y, float *z, double *sums) {
  PH_GUARD();
  Batch *B = (Batch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  if (((uintptr_t)x | (uintptr_t)y) % 16 != 0)
    return ph::set_error(PH_E_ARGUMENT, "batch: x and y must be 16-byte aligned");
  B->x = x;
  B->y = y;
  B->z = z;
  B->sums = sums;
  if (B->exec) {
    PH_HIP(hipGraphExecDestroy(B->exec));
    B->exec = nullptr;
  }
  if (B->graph) {
    PH_HIP(hipGraphDestroy(B->graph));
    B->graph = nullptr;
  }
  return PH_OK;
}

// mode 0: replay the captured graph (capturing it on first use)
// mode 1: launch the three kernels directly
// mode 2: per-array launches (three kernels per array)
int ph_batch_run(void *handle, int32_t mode) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  Batch *B = (Batch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  if (!B->x || !B->y || !B->z || !B->sums) return ph::set_error(PH_E_ARGUMENT, "batch buffers not bound");
  int rc = ph::launch_gate("batch");
  if (rc) return rc;
  if (B->n == 0) return PH_OK;
  if (mode == 0) {
    if (!B->exec) {
      rc = capture(B);
      if (rc) return rc;
    }
    PH_HIP(hipGraphLaunch(B->exec, ph::stream()));
    return PH_OK;
  }
  if (mode == 1) return launch_batch(B);
  if (mode == 2) {
    hipStream_t st = ph::stream();
    for (int k = 0; k < B->n; ++k) {
      const BArray &a = B->arrays[(size_t)k];
      const int64_t nz = (int64_t)a.zshp[0] * a.zshp[1] * a.zshp[2];
      const unsigned g1 = (unsigned)std::max<int64_t>(1, std::min<int64_t>((a.n + 255) / 256, 1024));
      const unsigned g3 = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nz + 255) / 256, 1024));
      hipLaunchKernelGGL(k_one_add, dim3(g1), dim3(256), 0, st, B->x, B->y, a.xoff, a.n, a.scalar);
      hipLaunchKernelGGL(k_one_sum, dim3(1), dim3(256), 0, st, B->x, a.xoff, a.n, B->sums + k);
      hipLaunchKernelGGL(k_one_slice, dim3(g3), dim3(256), 0, st, B->x, B->z, a, nz);
    }
    return ph::check_launch("per-array kernels");
  }
  return ph::set_error(PH_E_ARGUMENT, "batch mode %d", mode);
}

// Autotuner: times each candidate schedule (work-item size, workgroups per
// CU, queue or static assignment) on the bound buffers and keeps the fastest;
// the choice is cached per histogram of array sizes.  Writes the chosen work
// item size and grid to *chunk_out / *grid_out (may be NULL); dynamic_out
// (may be NULL) receives 1 for the queue schedule, 0 for the static one.
int ph_batch_tune(void *handle, int64_t *chunk_out, int32_t *grid_out, int32_t *dynamic_out) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  Batch *B = (Batch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  if (!B->x || !B->y || !B->z || !B->sums) return ph::set_error(PH_E_ARGUMENT, "batch buffers not bound");
  const std::string key = histogram_key(B);
  auto hit = tune_cache().find(key);
  Schedule best{16384, cus() * 4, 1};
  if (hit != tune_cache().end()) {
    best = hit->second;
  } else {
    hipEvent_t e0, e1;
    PH_HIP(hipEventCreate(&e0));
    PH_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    for (int64_t chunk : {1024, 2048, 4096, 8192, 32768}) {
      int rc = build_items(B, chunk);
      if (rc) return rc;
      for (int dyn : {1, 0}) {
        for (int per_cu : {2, 4, 8}) {
          B->grid = cus() * per_cu;
          B->dynamic = dyn;
          rc = launch_batch(B);  // warm
          if (rc) return rc;
          PH_HIP(hipEventRecord(e0, ph::stream()));
          for (int r = 0; r < 3; ++r) {
            rc = launch_batch(B);
            if (rc) return rc;
          }
          PH_HIP(hipEventRecord(e1, ph::stream()));
          PH_HIP(hipEventSynchronize(e1));
          float ms = 0;
          PH_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (ms < best_ms) {
            best_ms = ms;
            best = Schedule{chunk, B->grid, dyn};
          }
        }
      }
    }
    PH_HIP(hipEventDestroy(e0));
    PH_HIP(hipEventDestroy(e1));
    tune_cache()[key] = best;
  }
  int rc = build_items(B, best.chunk);
  if (rc) return rc;
  B->grid = best.grid;
  B->dynamic = best.dynamic;
  if (B->exec) {  // the schedule changed: recapture on next replay
    PH_HIP(hipGraphExecDestroy(B->exec));
    B->exec = nullptr;
    PH_HIP(hipGraphDestroy(B->graph));
    B->graph = nullptr;
  }
  if (chunk_out) *chunk_out = best.chunk;
  if (grid_out) *grid_out = best.grid;
  if (dynamic_out) *dynamic_out = best.dynamic;
  return PH_OK;
}

// Sets a schedule explicitly (chunk >= 64 elements, grid >= 1 workgroups).
int ph_batch_schedule(void *handle, int64_t chunk, int32_t grid, int32_t dynamic) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  Batch *B = (Batch *)handle;
  if (!B) return ph::set_error(PH_E_ARGUMENT, "null batch");
  if (chunk < 64 || chunk % 64 != 0 || chunk > (int64_t(1) << 30) || grid < 1 || grid > (1 << 20))
    return ph::set_error(PH_E_ARGUMENT, "batch schedule: chunk %lld, grid %d", (long long)chunk, grid);
  int rc = build_items(B, chunk);
  if (rc) return rc;
  B->grid = grid;
  B->dynamic = dynamic ? 1 : 0;
  if (B->exec) {
    PH_HIP(hipGraphExecDestroy(B->exec));
    B->exec = nullptr;
    PH_HIP(hipGraphDestroy(B->graph));
    B->graph = nullptr;
  }
  return PH_OK;
}

int ph_batch_destroy(void *handle) {
  PH_GUARD();
  Batch *B = (Batch *)handle;
  if (!B) return PH_OK;
  (void)hipStreamSynchronize(ph::stream());
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  (void)hipFree(B->d_recs);
  (void)hipFree(B->d_first);
  (void)hipFree(B->d_partial);
  (void)hipFree(B->d_counter);
  delete B;
  return PH_OK;
}

}  // extern "C"
