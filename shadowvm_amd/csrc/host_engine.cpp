// host_engine.cpp — compile, load and run the CPU engine's generated host code.
//
// The reference compiles each pipeline's host transform with
// `clang -O2 -fPIC -shared` into <home>/<id>.so and dlopens it by pipeline id
// (codegen/compiler.cpp:19-50).  Here the source comes from the same lowered plan as
// the kernels (gen_host_source); it is compiled by the host C++ compiler in a child
// process (posix_spawn), written to a private cache directory and loaded with dlopen.
// A cached object is used only when the key text stored beside it (compiler, options
// and the whole source) equals the request's.
#include "host_engine.h"

#include <dlfcn.h>
#include <fcntl.h>
#include <spawn.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>

extern char** environ;

namespace svm {

uint64_t source_hash(const std::string& s);

namespace {

std::mutex g_host_mu;
std::map<std::string, HostTaskFn> g_host_fns;  // source -> loaded task function

std::mutex g_hybrid_mu;
HybridConfig g_hybrid;

// The host compiler: ROCm's clang++ (the reference compiles its host code with clang),
// else the system g++.  -fwrapv: int32 overflow wraps, as in the kernels.
std::string host_compiler() {
  static const std::string cc = [] {
    const char* e = getenv("SVM_HOST_CXX");
    if (e && *e) return std::string(e);
    for (const char* p : {"/opt/rocm/llvm/bin/clang++", "/usr/bin/clang++", "/usr/bin/g++"})
      if (access(p, X_OK) == 0) return std::string(p);
    return std::string("g++");
  }();
  return cc;
}
const char* const kHostOpts[] = {"-O2", "-std=c++17", "-fwrapv", "-fPIC", "-shared", "-w"};

std::string private_dir() {
  const char* d = getenv("SVM_JIT_CACHE");
  std::string p = d && *d && std::string(d) != "off" ? d : "/tmp/svm_jit_cache-" + std::to_string((unsigned long)geteuid());
  mkdir(p.c_str(), 0700);
  struct stat st;
  if (lstat(p.c_str(), &st) != 0 || !S_ISDIR(st.st_mode) || st.st_uid != geteuid() ||
      (st.st_mode & (S_IWGRP | S_IWOTH)))
    fail(SVM_ERR_JIT, "host code directory " + p + " is not a private directory of this user");
  return p;
}

std::string read_all(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return {};
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

bool write_all(const std::string& path, const std::string& text) {
  std::ofstream f(path, std::ios::binary);
  f.write(text.data(), (std::streamsize)text.size());
  return (bool)f;
}

// Runs the compiler in a child process; returns its exit status and output.
int run_compiler(const std::vector<std::string>& args, std::string& log, const std::string& log_path) {
  std::vector<char*> argv;
  for (auto& a : args) argv.push_back(const_cast<char*>(a.c_str()));
  argv.push_back(nullptr);
  posix_spawn_file_actions_t fa;
  posix_spawn_file_actions_init(&fa);
  posix_spawn_file_actions_addopen(&fa, 1, log_path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0600);
  posix_spawn_file_actions_adddup2(&fa, 1, 2);
  pid_t pid = 0;
  const int rc = posix_spawn(&pid, argv[0], &fa, nullptr, argv.data(), environ);
  posix_spawn_file_actions_destroy(&fa);
  if (rc != 0) {
    log = std::string("cannot start ") + argv[0] + ": " + strerror(rc);
    return -1;
  }
  int status = 0;
  while (waitpid(pid, &status, 0) < 0 && errno == EINTR) {
  }
  log = read_all(log_path);
  unlink(log_path.c_str());
  return WIFEXITED(status) ? WEXITSTATUS(status) : -1;
}

HostTaskFn load(const std::string& so) {
  void* h = dlopen(so.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h) fail(SVM_ERR_JIT, std::string("dlopen of generated host code failed: ") + dlerror());
  void* f = dlsym(h, "svm_host_task");
  if (!f) fail(SVM_ERR_JIT, "generated host code has no svm_host_task");
  return (HostTaskFn)f;
}

}  // namespace

HostTaskFn host_task_function(const std::string& src) {
  std::lock_guard<std::mutex> g(g_host_mu);
  auto it = g_host_fns.find(src);
  if (it != g_host_fns.end()) return it->second;
  noalloc_refuse("to compile host code");
  std::string key = "svm-host-v1 " + host_compiler();
  for (const char* o : kHostOpts) key += std::string(" ") + o;
  key += "\n" + src;
  const std::string dir = private_dir();
  const std::string base = dir + "/host_" + std::to_string(source_hash(key));
  const std::string so = base + ".so", keyfile = base + ".key";
  if (read_all(keyfile) != key || access(so.c_str(), R_OK) != 0) {
    const std::string tag = std::to_string(getpid()) + "_" + std::to_string(g_host_fns.size());
    const std::string cpp = base + "_" + tag + ".cpp", tmp = base + "_" + tag + ".so";
    if (!write_all(cpp, src)) fail(SVM_ERR_JIT, "cannot write " + cpp);
    std::vector<std::string> args = {host_compiler()};
    for (const char* o : kHostOpts) args.push_back(o);
    args.push_back("-o");
    args.push_back(tmp);
    args.push_back(cpp);
    std::string log;
    const int rc = run_compiler(args, log, base + "_" + tag + ".log");
    unlink(cpp.c_str());
    if (rc != 0) {
      unlink(tmp.c_str());
      fail(SVM_ERR_JIT, "generated host code failed to compile:\n" + log + "\n--- source ---\n" + src);
    }
    if (rename(tmp.c_str(), so.c_str()) != 0) {
      unlink(tmp.c_str());
      fail(SVM_ERR_JIT, "cannot install " + so);
    }
    const std::string ktmp = keyfile + "." + tag;
    if (write_all(ktmp, key)) rename(ktmp.c_str(), keyfile.c_str());
  }
  HostTaskFn fn = load(so);
  g_host_fns[src] = fn;
  return fn;
}

HybridConfig hybrid_config() {
  std::lock_guard<std::mutex> g(g_hybrid_mu);
  return g_hybrid;
}

void set_hybrid_config(const HybridConfig& c) {
  if (c.cpu_ratio < 0 || c.cpu_ratio > 10) fail(SVM_ERR_INVALID, "cpu_ratio is 0..10 (bulks of every 10)");
  if (c.cpu_threads < 1 || c.cpu_threads > 1024) fail(SVM_ERR_INVALID, "cpu_threads is 1..1024");
  std::lock_guard<std::mutex> g(g_hybrid_mu);
  g_hybrid = c;
}

namespace {
struct Mirror {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;
  std::shared_ptr<const HostBytes> bytes;
};
std::mutex g_mirror_mu;
std::map<std::pair<const Table*, int>, Mirror> g_mirrors;
}  // namespace

std::shared_ptr<HostBytes> host_copy(const void* dev, uint64_t bytes, uint64_t pad) {
  noalloc_refuse("host copies of device data");
  auto v = std::make_shared<HostBytes>();
  v->data.reset(new char[bytes + pad + 1]);
  v->size = bytes + pad;
  if (bytes) SVM_HIP(hipMemcpy(v->data.get(), dev, bytes, hipMemcpyDeviceToHost));
  memset(v->data.get() + bytes, 0, pad + 1);
  return v;
}

std::shared_ptr<const HostBytes> host_mirror(const TablePtr& t, int col) {
  TablePtr b = t->base ? t->base : t;
  if (col < 0 || col >= (int)b->cols.size() || !b->cols[col].dev)
    fail(SVM_ERR_INVALID, "host mirror of a column that is not materialised");
  std::lock_guard<std::mutex> g(g_mirror_mu);
  auto key = std::make_pair(b.get(), col);
  auto it = g_mirrors.find(key);
  if (it != g_mirrors.end() && it->second.owner.lock() == b && it->second.version == b->data_version())
    return it->second.bytes;
  for (auto m = g_mirrors.begin(); m != g_mirrors.end();)
    m = m->second.owner.expired() ? g_mirrors.erase(m) : std::next(m);
  Mirror m;
  m.owner = b;
  m.version = b->data_version();
  m.bytes = host_copy(b->cols[col].dev, (b->row0 + b->rows) * (uint64_t)b->cols[col].size, kColumnPad);
  g_mirrors[key] = m;
  return m.bytes;
}

void run_host_tasks(size_t ntasks, int threads, const std::function<void(size_t)>& fn) {
  if (!ntasks) return;
  const size_t nth = std::min<size_t>(ntasks, (size_t)std::max(1, threads));
  std::atomic<size_t> next{0};
  std::mutex err_mu;
  std::exception_ptr err;
  auto body = [&] {
    for (;;) {
      const size_t t = next++;
      if (t >= ntasks) return;
      try {
        fn(t);
      } catch (...) {
        std::lock_guard<std::mutex> g(err_mu);
        if (!err) err = std::current_exception();
        next = ntasks;
        return;
      }
    }
  };
  std::vector<std::thread> pool;
  for (size_t k = 1; k < nth; ++k) pool.emplace_back(body);
  body();
  for (auto& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace svm
