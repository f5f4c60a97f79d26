// hipRTC module cache: compile once per (source, target), keep the module for
// the life of the process, and persist the code object on disk so a later
// process (or a test run) skips the compile.  The disk cache is keyed on the
// source, the target, the compile options and the hipRTC version.
#include "jit.hpp"

#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <vector>

#include "common.hpp"

namespace ph {

namespace {

struct Entry {
  hipModule_t module = nullptr;
};

std::map<std::string, Entry> &cache() {
  static std::map<std::string, Entry> c;
  return c;
}

long g_compiles = 0;

uint64_t fnv1a(const std::string &s) {
  uint64_t h = 1469598103934665603ULL;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ULL;
  }
  return h;
}

std::string cache_dir() {
  const char *d = std::getenv("PH_JIT_CACHE_DIR");
  if (d && *d) return d;
  const char *home = std::getenv("HOME");
  std::string base = (home && *home) ? std::string(home) + "/.cache" : std::string("/tmp");
  return base + "/phhip-jit";
}

void mkdirs(const std::string &path) {
  std::string cur;
  std::stringstream ss(path);
  std::string part;
  if (!path.empty() && path[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    mkdir(cur.c_str(), 0755);
  }
}

bool read_file(const std::string &path, std::vector<char> *out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !out->empty();
}

void write_file_atomic(const std::string &path, const std::vector<char> &data) {
  std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    f.write(data.data(), (std::streamsize)data.size());
  }
  std::rename(tmp.c_str(), path.c_str());
}

// Options of every hipRTC compile (besides the target): no contraction of
// a * b + c into one rounding, and no packed FP32 instructions
// (v_pk_add/mul/fma_f32) in generated code, exactly as for the
// ahead-of-time kernels.
const std::vector<std::string> &compile_options() {
  static const std::vector<std::string> o = {"-O3", "-ffp-contract=off", "-std=c++17", "-fno-fast-math",
                                             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  return o;
}

// Everything a code object depends on: the source, the target, the options
// and the hipRTC version.
std::string cache_key(const std::string &source, const std::string &arch) {
  int major = 0, minor = 0;
  hiprtcVersion(&major, &minor);
  std::string k = source + "\n// target " + arch + "\n// hiprtc " + std::to_string(major) + "." +
                  std::to_string(minor) + "\n// options";
  for (const std::string &o : compile_options()) k += " " + o;
  return k + "\n";
}

int compile(const std::string &source, const std::string &arch, std::vector<char> *code) {
  // PH_JIT_DUMP=<dir>: keep a copy of every generated source (for inspection)
  if (const char *dump = std::getenv("PH_JIT_DUMP")) {
    char name[64];
    std::snprintf(name, sizeof(name), "/%016llx.hip", (unsigned long long)fnv1a(source));
    mkdirs(dump);
    std::ofstream f(std::string(dump) + name);
    f << source;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source.c_str(), "ph_fused.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    return set_error(PH_E_HIP, "hiprtcCreateProgram failed");
  const std::string arch_opt = "--offload-arch=" + arch;
  std::vector<const char *> opts = {arch_opt.c_str()};
  for (const std::string &o : compile_options()) opts.push_back(o.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n + 1, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return set_error(PH_E_HIP, "hipRTC compile failed (%s): %.700s", hiprtcGetErrorString(r),
                     log.c_str());
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code->resize(n);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  ++g_compiles;
  if (const char *dump = std::getenv("PH_JIT_DUMP")) {
    char name[64];
    std::snprintf(name, sizeof(name), "/%016llx.co", (unsigned long long)fnv1a(source));
    std::ofstream f(std::string(dump) + name, std::ios::binary);
    f.write(code->data(), (std::streamsize)code->size());
  }
  return PH_OK;
}

}  // namespace

long jit_compiles() { return g_compiles; }

std::string jit_arch() {
  static std::string arch;
  if (arch.empty()) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device() < 0 ? 0 : device()) == hipSuccess) {
      arch = p.gcnArchName;
      size_t c = arch.find(':');
      if (c != std::string::npos) arch = arch.substr(0, c);
    }
    if (arch.empty()) arch = "gfx950";
  }
  return arch;
}

static int get_module(const std::string &source, hipModule_t *out) {
  auto &c = cache();
  auto it = c.find(source);
  if (it != c.end()) {
    *out = it->second.module;
    return PH_OK;
  }
  const std::string arch = jit_arch();
  const std::string key = cache_key(source, arch);
  char name[64];
  std::snprintf(name, sizeof(name), "%016llx-%s", (unsigned long long)fnv1a(key), arch.c_str());
  const std::string dir = cache_dir();
  const std::string path = dir + "/" + name + ".co", key_path = dir + "/" + name + ".key";
  std::vector<char> code, stored;
  hipModule_t mod = nullptr;
  // a cached object is used only when the key stored beside it is this one
  // (the file name is a 64-bit hash; the stored key rules out a collision)
  if (read_file(key_path, &stored) && std::string(stored.begin(), stored.end()) == key &&
      read_file(path, &code) && hipModuleLoadData(&mod, code.data()) == hipSuccess) {
    c[source].module = mod;
    *out = mod;
    return PH_OK;
  }
  (void)hipGetLastError();
  int rc = compile(source, arch, &code);
  if (rc) return rc;
  hipError_t e = hipModuleLoadData(&mod, code.data());
  if (e != hipSuccess)
    return set_error(PH_E_HIP, "hipModuleLoadData: %s", hipGetErrorString(e));
  mkdirs(dir);
  write_file_atomic(path, code);
  write_file_atomic(key_path, std::vector<char>(key.begin(), key.end()));
  c[source].module = mod;
  *out = mod;
  return PH_OK;
}

int jit_kernel(const std::string &source, const char *name, hipFunction_t *out) {
  hipModule_t mod;
  int rc = get_module(source, &mod);
  if (rc) return rc;
  hipError_t e = hipModuleGetFunction(out, mod, name);
  if (e != hipSuccess)
    return set_error(PH_E_HIP, "hipModuleGetFunction(%s): %s", name, hipGetErrorString(e));
  return PH_OK;
}

int jit_kernel2(const std::string &source, const char *name, hipFunction_t *out) {
  return jit_kernel(source, name, out);
}

int jit_compile_only(const std::string &source) {
  std::vector<char> code;
  return compile(source, jit_arch(), &code);
}

}  // namespace ph
