// plan.h — operator chain decoded from PipelineInfo, and expression lowering to HIP.
#pragma once

#include <set>
#include <string>
#include <vector>

#include "svm_internal.h"

namespace svm {

struct KeyIndex {
  int obj = 0;
  int col = 0;
};

// The operator chain of one pipeline, as Pipeline::Pipeline collects it
// (core/pipeline.cu:11-80): all Filter predicates (AND-ed in order), the last
// Map's select list, ReduceBy metric/dim lists, and JoinWith key pairs.
struct Plan {
  std::string id;
  bool has_source = false, has_sink = false, has_loader = false;
  std::string source_type;  // "object" | "file"
  std::string source_path;  // Source{type: file, path}
  std::string col_sep, row_sep;  // Loader{col_spilter, row_spilter}: their first characters are used
  ObjectInfo source_info, output_info, loader_info;
  std::vector<std::string> filter_exp, select_exp, metric_exp;
  std::vector<KeyIndex> reduce_key, join_lkey, join_rkey;
  bool is_reduce = false;
};

Plan decode_plan(const PipelineInfo& info);
std::vector<std::string> str_split(const std::string& code, const std::string& sep);
KeyIndex gen_key_index(const std::string& exp);

// Literal parameters collected while lowering (kernels are cached per shape,
// constants travel in the launch arguments).
struct Literals {
  std::vector<int32_t> ints;
  std::string blob;  // NUL-terminated string literals, 8-byte aligned entries
};

// Schemas of the pipeline's input objects: [0] = source, [k] = k-th ref object.
using Schemas = std::vector<std::vector<std::pair<int, int64_t>>>;

// How operands of a lowered expression are spelled in generated HIP code.
struct OperandNames {
  // obj 0 column c of Int32 type, already loaded for the current row
  std::string fact_reg(int c) const { return "f" + std::to_string(c); }
};

struct Lowered {
  std::string code;
  std::set<std::pair<int, int>> operands;  // (obj, col) referenced
};

// Lower one expression of the reference's language (C int expressions over
// $(o.c) operands, StrCmp / Hash calls, string literals), following
// CodeGen::GenExperssion (codegen/codegen.cpp:240-277): obj 0 reads the current
// row, obj k reads the joined dimension row; a String operand is a char* or, as
// a reduce key, Hash(ptr, size).  Decimal int32 literals become launch
// parameters PI(n); plain string literals become PS(offset).
Lowered lower_expr(const std::string& exp, const Schemas& schemas, bool is_reduce_key,
                   Literals* lits);

}  // namespace svm
