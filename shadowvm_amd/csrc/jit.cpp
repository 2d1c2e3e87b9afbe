// jit.cpp — hipRTC compilation of generated kernels for gfx950, cached per source.
//
// The reference writes generated CUDA to <home>/<id>.cu, runs `nvcc -O2 -ptx` in a
// child process and loads the PTX by pipeline id (codegen/compiler.cpp:52-84,99-109).
// Here the source is compiled in-process by hipRTC.  Code objects are cached per
// (source text, compile options, target, hipRTC version), so a pipeline *shape*
// compiles once whatever its constants and id: in memory, and as a file in a
// directory private to the user ($SVM_JIT_CACHE, default /tmp/svm_jit_cache-<uid>,
// created 0700 and used only when it is a directory owned by this user that no one
// else can write) so that later processes skip the compile.  A cache file holds the
// exact key text it was compiled from (options, target, version and the whole
// source) and is loaded only when that text equals the request's.
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
  std::string src;  // the source it was compiled from
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

// Compile options of every generated kernel.  -fwrapv: signed int32 overflow in the
// plan's expressions wraps (two's complement), as it does in practice in the
// reference's clang / nvcc builds; the CPU engine and the oracle compile with it too.
const char* const kOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-fwrapv", "-Xclang", "-target-feature",
                             "-Xclang", "-packed-fp32-ops"};
constexpr int kNumOpts = (int)(sizeof(kOpts) / sizeof(kOpts[0]));

// What a code object depends on besides the source: options, target, hipRTC version.
const std::string& key_prefix() {
  static const std::string p = [] {
    int major = 0, minor = 0;
    hiprtcVersion(&major, &minor);
    std::string k = "svm-jit-v2 hiprtc " + std::to_string(major) + "." + std::to_string(minor) + " opts";
    for (const char* o : kOpts) k += std::string(" ") + o;
    return k + "\n";
  }();
  return p;
}

// The cache directory, or "" when there is none this process may trust (read on
// every compile, so SVM_JIT_CACHE may change between runs; "off" disables it).
std::string cache_dir() {
  const char* d = getenv("SVM_JIT_CACHE");
  std::string p = d && *d ? d : "/tmp/svm_jit_cache-" + std::to_string((unsigned long)geteuid());
  if (p == "off") return std::string();
  mkdir(p.c_str(), 0700);
  struct stat st;
  if (lstat(p.c_str(), &st) != 0 || !S_ISDIR(st.st_mode) || st.st_uid != geteuid() ||
      (st.st_mode & (S_IWGRP | S_IWOTH))) {
    static std::atomic<bool> warned{false};
    if (!warned.exchange(true))
      fprintf(stderr, "[svm] JIT cache directory %s is not a private directory of this user; "
                      "code objects are not cached on disk\n", p.c_str());
    return std::string();
  }
  return p;
}

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return {};
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// Cache file layout: "SVMCO2\n", the key length (8 bytes), the key text, then the
// code object.  The key text is key_prefix() + the source.
const char kMagic[] = "SVMCO2\n";

std::string cached_code(const std::string& path, const std::string& key) {
  std::string f = read_file(path);
  const size_t m = sizeof(kMagic) - 1;
  if (f.size() < m + 8 || f.compare(0, m, kMagic) != 0) return {};
  uint64_t n = 0;
  memcpy(&n, f.data() + m, 8);
  if (n != key.size() || f.size() <= m + 8 + n || f.compare(m + 8, n, key) != 0) return {};
  return f.substr(m + 8 + n);
}

void store_code(const std::string& path, const std::string& key, const std::string& code) {
  std::string tmp = path + ".tmp" + std::to_string(getpid()) + "." + std::to_string(g_compiles.load());
  {
    std::ofstream f(tmp, std::ios::binary);
    const uint64_t n = key.size();
    f.write(kMagic, sizeof(kMagic) - 1);
    f.write((const char*)&n, 8);
    f.write(key.data(), (std::streamsize)key.size());
    f.write(code.data(), (std::streamsize)code.size());
    if (!f) {
      f.close();
      unlink(tmp.c_str());
      return;
    }
  }
  if (rename(tmp.c_str(), path.c_str()) != 0) unlink(tmp.c_str());
}

std::string compile(const std::string& src) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "svm_kernel.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    fail(SVM_ERR_JIT, "hiprtcCreateProgram failed");
  hiprtcResult r = hiprtcCompileProgram(prog, kNumOpts, kOpts);
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
  return code;
}

}  // namespace

std::string jit_code_object(const std::string& src) {
  const std::string key = key_prefix() + src;
  const uint64_t h = source_hash(key);
  // SVM_JIT_DUMP=<dir>: keep every generated source (for offline inspection)
  const char* dump = getenv("SVM_JIT_DUMP");
  if (dump && *dump) {
    mkdir(dump, 0700);
    std::ofstream f(std::string(dump) + "/" + std::to_string(h) + ".hip", std::ios::binary);
    f << src;
  }
  const std::string dir = cache_dir();
  const std::string path = dir.empty() ? std::string() : dir + "/" + std::to_string(h) + ".co";
  std::string code = path.empty() ? std::string() : cached_code(path, key);
  if (code.empty()) {
    code = compile(src);
    if (!path.empty()) store_code(path, key, code);
  }
  return code;
}

hipFunction_t jit_function(const std::string& src, const char* name) {
  return jit_function(src, source_hash(src), name);
}

hipFunction_t jit_function(const std::string& src, uint64_t h, const char* name) {
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_mods.find(h);
  if (it != g_mods.end() && it->second.src != src) it = g_mods.end();  // a 64-bit hash collision
  if (it == g_mods.end()) {
    noalloc_refuse("to compile and load a kernel not loaded yet");
    std::string code = jit_code_object(src);
    Module m;
    m.src = src;
    hipError_t e = hipModuleLoadData(&m.mod, code.data());
    if (e != hipSuccess) {
      // a cache file that does not load on this device: compile afresh
      code = compile(src);
      SVM_HIP(hipModuleLoadData(&m.mod, code.data()));
    }
    auto old = g_mods.find(h);
    if (old != g_mods.end()) g_mods.erase(old);  // the colliding module stays loaded (its functions may be in use)
    it = g_mods.emplace(h, std::move(m)).first;
  }
  auto f = it->second.fns.find(name);
  if (f != it->second.fns.end()) return f->second;
  noalloc_refuse("to look up a kernel function not used yet");
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
