// Run-time kernel generation for fused elementwise programs (hipRTC).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace ph {

// Compiles `source` (once per process; also cached on disk under
// $PH_JIT_CACHE_DIR, default ~/.cache/phhip-jit, keyed by the source, the
// target, the compile options and the hipRTC version) and returns the kernel named `name`.
int jit_kernel(const std::string &source, const char *name, hipFunction_t *out);

// Same module, another kernel of it (the module must already be compiled).
int jit_kernel2(const std::string &source, const char *name, hipFunction_t *out);

// Compiles without loading (no device needed).
int jit_compile_only(const std::string &source);

// The device's target name (e.g. "gfx950").
std::string jit_arch();

// Number of modules compiled by hipRTC (not loaded from a cache) so far.
long jit_compiles();

}  // namespace ph
