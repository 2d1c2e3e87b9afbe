// jit.cpp — hipRTC compilation of generated kernels for gfx950, cached per source.
//
// The reference writes generated CUDA to <home>/<id>.cu, runs `nvcc -O2 -ptx` in a
// child process and loads the PTX by pipeline id (codegen/compiler.cpp:52-84,99-109).
// Here the source is compiled in-process by hipRTC, the code object is cached by a
// hash of the source text (so a pipeline *shape* compiles once, whatever its
// constants and id), both in memory and as a file under $SVM_JIT_CACHE
// (default /tmp/svm_jit_cache) so later processes skip the compile.
#include "jit.h"

#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace svm {

namespace {

struct Module {
  hipModule_t mod = nullptr;
  std::map<std::string, hipFunction_t> fns;
};

std::mutex g_mu;
std::map<uint64_t, Module> g_mods;
std::atomic<uint64_t> g_compiles{0};


}  // namespace

// 64-bit multiply-xorshift hash, eight bytes per step (a kernel source is hashed on
// every run of a pipeline, so this stays a few microseconds for 10 KB).
uint64_t source_hash(const std::string& s) {
  const uint64_t k = 0x9E3779B97F4A7C15ull;
  const size_t n = s.size();
  uint64_t h = 0xCBF29CE484222325ull ^ ((uint64_t)n * k);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, s.data() + i, 8);
    h = (h ^ w) * k;
    h ^= h >> 32;
  }
  uint64_t w = 0;
  memcpy(&w, s.data() + i, n - i);
  h = (h ^ w) * k;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  return h;
}

namespace {

std::string cache_dir() {
  const char* d = getenv("SVM_JIT_CACHE");
  std::string dir = d && *d ? d : "/tmp/svm_jit_cache";
  mkdir(dir.c_str(), 0777);
  return dir;
}

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return {};
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::string compile(const std::string& src, uint64_t h) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "svm_kernel.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    fail(SVM_ERR_JIT, "hiprtcCreateProgram failed");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    fail(SVM_ERR_JIT, "generated kernel failed to compile:\n" + log + "\n--- source ---\n" + src);
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  std::string code(n, '\0');
  hiprtcGetCode(prog, &code[0]);
  hiprtcDestroyProgram(&prog);
  g_compiles++;
  std::string path = cache_dir() + "/" + std::to_string(h) + ".co";
  std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    f.write(code.data(), (std::streamsize)code.size());
  }
  rename(tmp.c_str(), path.c_str());
  return code;
}

}  // namespace

std::string jit_code_object(const std::string& src) {
  uint64_t h = source_hash(src);
  // SVM_JIT_DUMP=<dir>: keep every generated source (for offline inspection)
  const char* dump = getenv("SVM_JIT_DUMP");
  if (dump && *dump) {
    mkdir(dump, 0777);
    std::ofstream f(std::string(dump) + "/" + std::to_string(h) + ".hip", std::ios::binary);
    f << src;
  }
  std::string path = cache_dir() + "/" + std::to_string(h) + ".co";
  std::string code = read_file(path);
  if (code.empty()) code = compile(src, h);
  return code;
}

hipFunction_t jit_function(const std::string& src, const char* name) {
  uint64_t h = source_hash(src);
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_mods.find(h);
  if (it == g_mods.end()) {
    std::string code = jit_code_object(src);
    Module m;
    hipError_t e = hipModuleLoadData(&m.mod, code.data());
    if (e != hipSuccess) {
      // A stale or foreign cache file: recompile once.
      code = compile(src, h);
      SVM_HIP(hipModuleLoadData(&m.mod, code.data()));
    }
    it = g_mods.emplace(h, std::move(m)).first;
  }
  auto f = it->second.fns.find(name);
  if (f != it->second.fns.end()) return f->second;
  hipFunction_t fn = nullptr;
  SVM_HIP(hipModuleGetFunction(&fn, it->second.mod, name));
  it->second.fns[name] = fn;
  return fn;
}

uint64_t jit_compile_count() { return g_compiles.load(); }

void launch(hipFunction_t fn, unsigned grid, unsigned block, unsigned lds, hipStream_t s,
            const SvmArgs& args) {
  SvmArgs copy = args;
  void* params[] = {&copy};
  SVM_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, lds, s, params, nullptr));
}

}  // namespace svm
