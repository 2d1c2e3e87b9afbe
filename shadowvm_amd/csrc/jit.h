// jit.h — hipRTC kernel cache and launch helper.
#pragma once

#include <cstdint>
#include <string>

#include "kargs.h"
#include "svm_internal.h"

namespace svm {

// Compile (or fetch from the cache) the module built from `src` and return the
// named kernel.  Throws SVM_ERR_JIT with the compiler log on failure.
hipFunction_t jit_function(const std::string& src, const char* name);
// The same for a source whose source_hash is already known.
hipFunction_t jit_function(const std::string& src, uint64_t hash, const char* name);
// Code object for `src` (compiled if not cached); used to warm the cache.
std::string jit_code_object(const std::string& src);
uint64_t jit_compile_count();
// The hash a kernel source is cached under.
uint64_t source_hash(const std::string& s);
void launch(hipFunction_t fn, unsigned grid, unsigned block, unsigned lds, hipStream_t s,
            const SvmArgs& args);

}  // namespace svm
