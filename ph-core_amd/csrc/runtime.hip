// Runtime of libphhip: device selection, the library stream, device buffers,
// copies, events, error state, descriptor validation.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace ph {

namespace {
thread_local char g_err[1024] = "";
bool g_inited = false;
int g_device = -1;
hipStream_t g_stream = nullptr;
struct Scratch {
  void *ptr = nullptr;
  size_t bytes = 0;
};
Scratch g_scratch[12];
long g_launches = 0;
long g_fail_after = -2;  // -2: not read yet, -1: disabled
}  // namespace

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

std::recursive_mutex &abi_mutex() {
  static std::recursive_mutex m;
  return m;
}

bool inited() { return g_inited; }
hipStream_t stream() { return g_stream; }
int device() { return g_device; }

int scratch(size_t bytes, void **out, int slot) {
  Scratch &s = g_scratch[slot];
  if (bytes > s.bytes) {
    if (s.ptr) {
      PH_HIP(hipStreamSynchronize(g_stream));
      PH_HIP(hipFree(s.ptr));
      s.ptr = nullptr;
      s.bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (hipMalloc(&s.ptr, want) != hipSuccess) {
      s.ptr = nullptr;
      return set_error(PH_E_OOM, "scratch allocation of %zu bytes failed", want);
    }
    s.bytes = want;
  }
  *out = s.ptr;
  return PH_OK;
}

size_t dtype_size(int dt) {
  switch (dt) {
    case PH_I32: return 4;
    case PH_F32: return 4;
    case PH_F64: return 8;
    case PH_BF16: return 2;
    case PH_BOOL: return 1;
  }
  return 0;
}

int64_t desc_count(const ph_view_desc *d) {
  int64_t n = 1;
  for (int i = 0; i < d->view.rank; ++i) n *= d->view.shape[i];
  return n;
}

// [lo, hi] of offset + sum c_i * stride_i over the box c_i in [0, shape_i).
static void affine_range(const ph_affine &a, int64_t *lo, int64_t *hi) {
  int64_t l = a.offset, h = a.offset;
  for (int i = 0; i < a.rank; ++i) {
    int64_t span = (a.shape[i] - 1) * a.stride[i];
    if (span < 0) l += span;
    else h += span;
  }
  *lo = l;
  *hi = h;
}

int validate_desc(const ph_view_desc *d, const char *what) {
  if (!d) return set_error(PH_E_ARGUMENT, "%s: null descriptor", what);
  if (dtype_size(d->dtype) == 0)
    return set_error(PH_E_ARGUMENT, "%s: unknown dtype %d", what, d->dtype);
  if (d->view.rank < 0 || d->view.rank > PH_MAX_RANK)
    return set_error(PH_E_DIMENSION, "%s: rank %d outside [0, %d]", what, d->view.rank, PH_MAX_RANK);
  if (d->nstage < 0 || d->nstage > PH_MAX_STAGES)
    return set_error(PH_E_ARGUMENT, "%s: %d stages (max %d)", what, d->nstage, PH_MAX_STAGES);
  for (int i = 0; i < d->view.rank; ++i)
    if (d->view.shape[i] < 0)
      return set_error(PH_E_DIMENSION, "%s: negative extent on axis %d", what, i);
  if (desc_count(d) == 0) return PH_OK;  // addresses nothing
  if (d->extent <= 0)
    return set_error(PH_E_SHAPE, "%s: view of %lld elements over an empty buffer", what,
                     (long long)desc_count(d));
  int64_t lo, hi;
  affine_range(d->view, &lo, &hi);
  for (int s = 0; s < d->nstage; ++s) {
    const ph_affine &st = d->stage[s];
    if (st.rank < 1 || st.rank > PH_MAX_RANK)
      return set_error(PH_E_DIMENSION, "%s: stage %d rank %d", what, s, st.rank);
    int64_t cells = 1;
    for (int j = 0; j < st.rank; ++j) {
      if (st.shape[j] <= 0)
        return set_error(PH_E_DIMENSION, "%s: stage %d extent %lld", what, s, (long long)st.shape[j]);
      cells *= st.shape[j];
    }
    if (lo < 0 || hi >= cells)
      return set_error(PH_E_INDEX, "%s: stage %d input index range [%lld, %lld] outside [0, %lld)",
                       what, s, (long long)lo, (long long)hi, (long long)cells);
    affine_range(st, &lo, &hi);
  }
  if (lo < 0 || hi >= d->extent)
    return set_error(PH_E_INDEX, "%s: addresses [%lld, %lld] outside the buffer [0, %lld)", what,
                     (long long)lo, (long long)hi, (long long)d->extent);
  return PH_OK;
}

void collapse(const ph_view_desc *const *descs, int nops, Layout *out) {
  const int rank = descs[0]->view.rank;
  out->nops = nops;
  // drop extent-1 dimensions
  int keep[PH_MAX_RANK];
  int nk = 0;
  for (int i = 0; i < rank; ++i)
    if (descs[0]->view.shape[i] != 1) keep[nk++] = i;
  int64_t shape[PH_MAX_RANK];
  int64_t stride[kMaxOps][PH_MAX_RANK];
  for (int k = 0; k < nk; ++k) {
    shape[k] = descs[0]->view.shape[keep[k]];
    for (int o = 0; o < nops; ++o) stride[o][k] = descs[o]->view.stride[keep[k]];
  }
  for (int o = 0; o < nops; ++o) out->offset[o] = descs[o]->view.offset;
  // merge k into k+1 (outer into inner) where every operand is contiguous
  int r = 0;
  for (int k = 0; k < nk; ++k) {
    if (r > 0) {
      bool ok = true;
      for (int o = 0; o < nops; ++o)
        if (out->stride[o][r - 1] != stride[o][k] * shape[k]) ok = false;
      if (ok) {
        out->shape[r - 1] *= shape[k];
        for (int o = 0; o < nops; ++o) out->stride[o][r - 1] = stride[o][k];
        continue;
      }
    }
    out->shape[r] = shape[k];
    for (int o = 0; o < nops; ++o) out->stride[o][r] = stride[o][k];
    ++r;
  }
  out->rank = r;
}

long launch_count() { return g_launches; }

int launch_gate(const char *what) {
  if (g_fail_after == -2) {
    const char *e = std::getenv("PH_HIP_FAIL_AFTER");
    g_fail_after = e ? std::atol(e) : -1;
  }
  if (g_fail_after >= 0 && g_launches >= g_fail_after)
    return set_error(PH_E_HIP, "%s: launch refused by PH_HIP_FAIL_AFTER=%ld", what, g_fail_after);
  ++g_launches;
  return PH_OK;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(PH_E_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
  return PH_OK;
}

}  // namespace ph

extern "C" {

long ph_launch_count(void) { return ph::launch_count(); }


int ph_init(int device) {
  PH_GUARD();
  if (ph::g_inited) {
    if (device == ph::g_device) return PH_OK;
    return ph::set_error(PH_E_ARGUMENT, "already initialised on device %d", ph::g_device);
  }
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return ph::set_error(PH_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n)
    return ph::set_error(PH_E_ARGUMENT, "device %d outside [0, %d)", device, n);
  PH_HIP(hipSetDevice(device));
  PH_HIP(hipStreamCreateWithFlags(&ph::g_stream, hipStreamNonBlocking));
  const int rc = ph::heat_noalloc_setup(device);
  if (rc) {
    (void)hipStreamDestroy(ph::g_stream);
    ph::g_stream = nullptr;
    return rc;
  }
  ph::g_device = device;
  ph::g_inited = true;
  return PH_OK;
}

int ph_device_count(int *n) {
  PH_GUARD();
  if (!n) return ph::set_error(PH_E_ARGUMENT, "ph_device_count: null output");
  *n = 0;
  hipError_t e = hipGetDeviceCount(n);
  if (e != hipSuccess) {
    *n = 0;
    return ph::set_error(PH_E_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return PH_OK;
}

const char *ph_last_error(void) { return ph::g_err; }

int ph_alloc(void **ptr, size_t bytes) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  *ptr = nullptr;
  if (bytes == 0) bytes = 256;  // a valid, distinct pointer for empty arrays
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return ph::set_error(PH_E_OOM, "hipMalloc of %zu bytes failed", bytes);
  }
  return PH_OK;
}

int ph_free(void *ptr) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (!ptr) return PH_OK;
  PH_HIP(hipStreamSynchronize(ph::g_stream));
  PH_HIP(hipFree(ptr));
  return PH_OK;
}

int ph_h2d(void *dst, const void *src, size_t bytes) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (bytes == 0) return PH_OK;
  PH_HIP(hipStreamSynchronize(ph::g_stream));
  PH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ph::g_stream));
  PH_HIP(hipStreamSynchronize(ph::g_stream));
  return PH_OK;
}

int ph_d2h(void *dst, const void *src, size_t bytes) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (bytes == 0) return PH_OK;
  PH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ph::g_stream));
  PH_HIP(hipStreamSynchronize(ph::g_stream));
  return PH_OK;
}

int ph_d2d(void *dst, const void *src, size_t bytes) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  if (bytes == 0) return PH_OK;
  PH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ph::g_stream));
  return PH_OK;
}

int ph_sync(void) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  PH_HIP(hipStreamSynchronize(ph::g_stream));
  return PH_OK;
}

int ph_get_stream(void **s) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  *s = (void *)ph::g_stream;
  return PH_OK;
}

int ph_event_create(void **ev) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  hipEvent_t e;
  PH_HIP(hipEventCreate(&e));
  *ev = (void *)e;
  return PH_OK;
}

int ph_event_record(void *ev) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  PH_HIP(hipEventRecord((hipEvent_t)ev, ph::g_stream));
  return PH_OK;
}

int ph_event_elapsed(float *ms, void *start, void *stop) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  PH_HIP(hipEventSynchronize((hipEvent_t)stop));
  PH_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return PH_OK;
}

int ph_event_destroy(void *ev) {
  PH_GUARD();
  PH_REQUIRE_INIT();
  PH_HIP(hipEventDestroy((hipEvent_t)ev));
  return PH_OK;
}

}  // extern "C"
