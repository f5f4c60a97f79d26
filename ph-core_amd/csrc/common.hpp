// Shared host-side helpers of libphhip: error state, the library stream,
// descriptor validation and dimension collapsing, launch bookkeeping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <mutex>

#include "../../include/phhip.h"

namespace ph {

// Sets the thread's last-error text and returns `code`.
int set_error(int code, const char *fmt, ...);

// Serialises the ABI (ph-core's objects are single-threaded; the C side keeps
// one stream and one set of scratch buffers).
std::recursive_mutex &abi_mutex();

bool inited();
hipStream_t stream();
int device();

// Lazily grown device scratch, valid until the next call that asks for more;
// all work on it is ordered by the single library stream.  Slots: 0-1
// gathered map and reduction inputs, 2 reduction and heat partials, 3 flag
// words, 4-7 the buffers of maps whose inputs share the output's buffer,
// 8 the inverse of an indexed map's write index, 9-11 the marks, depths and
// level counts of an ordered map's tree of dependents.
int scratch(size_t bytes, void **out, int slot = 0);

// ph_strided_copy when source and destination share a buffer: the copy in
// the reference's lexicographic order (each element read after the writes
// of the elements before it), run in parallel where that order cannot be
// observed.  *done is false when no element reads another's destination,
// and the ordinary copy kernels then give the same result.
int ordered_copy(void *dst, const ph_view_desc *dd, const void *src, const ph_view_desc *sd, bool *done);

size_t dtype_size(int dt);

// Number of elements of the descriptor's logical shape (0 if any extent is 0).
int64_t desc_count(const ph_view_desc *d);

// Rejects malformed descriptors and any that can address outside
// [0, d->extent).  `what` names the operand in the error text.
int validate_desc(const ph_view_desc *d, const char *what);

// Layout of several operands over one logical shape after merging adjacent
// dimensions that are contiguous for every operand and dropping extent-1
// dimensions.  rank 0 means a single element.
constexpr int kMaxOps = PH_MAX_INPUTS + 1;
struct Layout {
  int rank;
  int nops;
  int64_t shape[PH_MAX_RANK];
  int64_t offset[kMaxOps];
  int64_t stride[kMaxOps][PH_MAX_RANK];
};
// All descriptors must be affine (nstage == 0) with equal logical shapes.
void collapse(const ph_view_desc *const *descs, int nops, Layout *out);

// Counts launches; PH_HIP_FAIL_AFTER=n makes launch n+1 (and later) fail with
// PH_E_HIP before it is issued, so host code can exercise its device-error
// handling without a faulting kernel.
int launch_gate(const char *what);
long launch_count();

// Generates and compiles (no launch) the general batch kernels for `prog`
// as map and, when its result is a float, as reduction program, over
// elements of x_dtype (ph_compile_check kind 5).
int gbatch_compile_check(const ph_prog *prog, int x_dtype);

// Looks up, for `device` (current), the device address of the fixed array
// ph_heat2d_noalloc keeps its partial sums in (heat.hip).
int heat_noalloc_setup(int device);
int check_launch(const char *what);

// Zeroes n int32 words on the library stream with a kernel (a memset node
// would misbehave inside a replayed HIP graph).
int clear_i32(int *p, int n);

}  // namespace ph

#define PH_GUARD() std::lock_guard<std::recursive_mutex> ph_guard_(ph::abi_mutex())

#define PH_REQUIRE_INIT()                                                    \
  do {                                                                       \
    if (!ph::inited())                                                       \
      return ph::set_error(PH_E_NOINIT, "ph_init has not been called");     \
  } while (0)

#define PH_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ph::set_error(PH_E_HIP, "%s failed: %s", #call,                 \
                           hipGetErrorString(e_));                           \
  } while (0)
