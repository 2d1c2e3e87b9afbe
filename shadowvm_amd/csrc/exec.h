// exec.h — executing one decoded pipeline on the GPU.
#pragma once

#include <memory>

#include "plan.h"
#include "svm_internal.h"

// Opaque C-ABI result (include/svm.h).
struct svm_result {
  int kind = 0;  // 0 select, 1 group-by, 2 scalar
  int64_t rows = 0;
  std::vector<std::pair<int, int64_t>> types;
  std::vector<std::vector<uint8_t>> cols;
  double build_ms = 0, scan_ms = 0, finish_ms = 0;
  std::string tune;  // launch configuration chosen for the scan
};

namespace svm {

struct RunOptions {
  bool want_host = true;        // copy the output back to the host
  hipStream_t stream = nullptr; // null: the engine stream
  int block = 256;
  int unroll = 1;
  int blocks_per_cu = 0;        // 0: derived from LDS use
  // No-allocation run of a fact (ReduceBy) pipeline, inside a NoAllocScope: the
  // group keys and sums are written into these caller device buffers (K Int32 key
  // columns then M Int64 sum columns, out_cap rows each), no output object is
  // created or cached, no autotuning trial runs, and the run is untimed.  *out_rows
  // receives the group count (also when it exceeds out_cap, which is an error).
  bool noalloc = false;
  void* const* out_dev = nullptr;
  int out_ncols = 0;
  uint64_t out_cap = 0;
  uint64_t* out_rows = nullptr;
};

// Runs `info` against the store (pre_pipeline_id resolution as RunPipelineCMD,
// service/service.cu:32-52) and caches the output table under info.id.
std::unique_ptr<svm_result> run_pipeline(const PipelineInfo& info, const RunOptions& opt);

TablePtr gen_device_table(const std::string& id, int kind, uint64_t rows, uint64_t seed,
                          const svm_gen_params* p);
void gen_host(int kind, uint64_t seed, const svm_gen_params* p, uint64_t row0, uint64_t n,
              void* const* cols, int threads);
std::vector<std::pair<int, int64_t>> gen_schema(int kind);
void engine_init(const Config& cfg);
void engine_shutdown();

}  // namespace svm
