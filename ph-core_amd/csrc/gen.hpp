// Lowering of element programs (ph_prog) to HIP source, shared by the fused
// map and reduction kernels (fuse.cpp) and the general batch (gbatch.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/phhip.h"

namespace ph {
namespace gen {

const char *ctype(int dt);  // compute type of a dtype (BF16 computes as float)
const char *stype(int dt);  // storage type
int compute_dt(int storage);
std::string fmt(const char *f, ...);
std::string lit_i64(int64_t v);
bool is_float(int dt);
// Checks `p` against the input storage dtypes (PH_OP_INDEX only where
// `allow_index` says the kernel defines it).
int validate_prog(const ph_prog *p, int n_in, const int *in_dt, bool allow_index = false);
bool prog_uses_index(const ph_prog *p);
// Whether any element can raise a PH_FLAG_* condition.
bool prog_may_flag(const ph_prog *p);
// Appends the statements computing `p` over the input expressions `inexpr`
// (of their compute types) to *code, naming values with suffix `sfx`;
// returns the result's variable name.  Statements set `fl` (an int in scope)
// for flagged conditions.
std::string emit_prog(const ph_prog *p, const std::vector<std::string> &inexpr, std::string *code,
                      const std::string &sfx);
// Loads, stores and arithmetic helpers every generated kernel starts with.
extern const char *const kPrelude;

}  // namespace gen
}  // namespace ph
