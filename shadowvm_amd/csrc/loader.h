// loader.h — the Loader operator: '|'-separated text files parsed into a device
// table (core/operator.cpp:13-48, object/object.cpp:122-169, com/cpp/com.cu:33-44).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>

#include "exec.h"
#include "plan.h"

namespace svm {
namespace ld {

constexpr int kMaxLoadCols = 24;  // MaxColumnNum, com/cpp/define.hpp:26

struct ParseArgs {
  const char* buf;              // the file's bytes in HBM
  uint64_t n;                   // bytes
  const unsigned long long* ends;  // position of each row separator
  uint64_t rows;                // rows of this file
  uint64_t row0;                // first output row of this file
  char col_sep;
  int ncols;
  int type[kMaxLoadCols];
  int size[kMaxLoadCols];
  void* col[kMaxLoadCols];      // output columns (row 0)
};

// Separators per 16 KB chunk of buf[0..n) (buf padded with 64 readable bytes).
void count_separators(const char* buf, uint64_t n, char sep, unsigned* counts, uint64_t nchunks, hipStream_t s);
// ends[offsets[c] + k] = position of the k-th separator of chunk c.
void separator_positions(const char* buf, uint64_t n, char sep, const unsigned long long* offsets,
                         unsigned long long* ends, uint64_t nchunks, hipStream_t s);
void parse_rows(const ParseArgs& a, hipStream_t s);
constexpr uint64_t kChunk = 16384;

}  // namespace ld

// Runs a Loader pipeline: Source{type: file, path} + Loader{col_spilter,
// row_spilter} + Sink; the table is cached under the pipeline id.
std::unique_ptr<svm_result> run_load(const PipelineInfo& info, const Plan& plan, const RunOptions& opt);

}  // namespace svm
