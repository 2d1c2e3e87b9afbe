// batch.cpp — batched execution and the serving loop.
#include "exec.h"
#include "svm_internal.h"

namespace svm {

int batch_run(int n, const uint8_t* const* plans, const size_t* lens, svm_result** outs) {
  if (n < 0 || (n && (!plans || !lens))) fail(SVM_ERR_INVALID, "bad batch arguments");
  Engine& E = engine();
  std::lock_guard<std::mutex> g(E.run_mu);
  for (int q = 0; q < n; ++q) {
    PipelineInfo info = decode_pipeline_info(plans[q], lens[q]);
    RunOptions opt;
    opt.want_host = outs != nullptr;
    auto r = run_pipeline(info, opt);
    if (outs) outs[q] = r.release();
  }
  return SVM_OK;
}

int server_start(int, uint64_t) { fail(SVM_ERR_UNSUPPORTED, "serving loop not built yet"); }
int server_submit(const uint8_t*, size_t, uint64_t*) { fail(SVM_ERR_STATE, "server not running"); }
int server_wait(uint64_t, int, svm_result**) { fail(SVM_ERR_STATE, "server not running"); }
int server_stop() { return SVM_OK; }

}  // namespace svm
