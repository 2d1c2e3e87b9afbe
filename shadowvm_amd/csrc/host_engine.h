// host_engine.h — the CPU engine: a pipeline's per-row transform run on host cores,
// for PipelineInfo.isGPU = false and for the host share of hybrid execution.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "kargs.h"
#include "svm_internal.h"

namespace svm {

// A compiled host task function (see gen_host_source).
using HostTaskFn = void (*)(const SvmArgs*, long long, long long, SvmHostOut*);
// Compiles `src` with the host C++ compiler into a shared object and loads it, as
// JitCompiler::CompileHostFuc does (codegen/compiler.cpp:19-50); cached per source
// in memory and, verified against the full source, on disk next to the kernel cache.
HostTaskFn host_task_function(const std::string& src);

// Process-wide hybrid settings (ResMgr::enable_hybrid_exec_, CPU_ratio_ and CPU_num_,
// core/resource.cu:14-21): with hybrid on, bulk i of a GPU pipeline runs on the host
// when i % 10 < cpu_ratio; at most cpu_threads host tasks run at once.
struct HybridConfig {
  bool enabled = false;
  int cpu_ratio = 5;
  int cpu_threads = 16;
};
HybridConfig hybrid_config();
void set_hybrid_config(const HybridConfig& c);

// Host memory holding a copy of device bytes (not zero-initialised beyond the copy
// and its zero padding).
struct HostBytes {
  std::unique_ptr<char[]> data;
  uint64_t size = 0;
  const char* ptr() const { return data.get(); }
};

// The bytes of column `col` of table `t` (its base table for a view) on the host,
// indexed like the device column (row r of the base at r * size), plus
// kColumnPad zero bytes.  Copied from HBM once per table contents and cached.
std::shared_ptr<const HostBytes> host_mirror(const TablePtr& t, int col);

// `bytes` bytes of device memory copied to a new host buffer (plus `pad` zero bytes).
std::shared_ptr<HostBytes> host_copy(const void* dev, uint64_t bytes, uint64_t pad = 0);

// Runs fn(task) for task in [0, ntasks) on at most `threads` host threads (the CPU
// task slots of ResMgr::GetCPU, core/resource.cu:89-105); rethrows the first error.
void run_host_tasks(size_t ntasks, int threads, const std::function<void(size_t)>& fn);

}  // namespace svm
