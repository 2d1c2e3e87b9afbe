// scratch.hip — the device scratch region of svm_pipeline_run_noalloc.
//
// That entry point allocates nothing: the per-run temporaries of a fact pipeline
// (dense join maps, group tables, emission offsets, literal blobs), which
// svm_pipeline_run takes from a per-stream arena grown with hipMalloc, come from this
// file-scope device array instead.  Each GPU holds its own copy of the array, at its
// own address, for the life of the process.  In host code `scratch` is only a
// host-side stand-in for the symbol: its device address is looked up once per device
// with hipGetSymbolAddress (svm_init calls scratch_init with the device current), and
// only that address is ever handed to kernels or copies.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "svm_internal.h"

namespace svm {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;  // device ordinal -> device address of `scratch`

}  // namespace

void scratch_init(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    fail(SVM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  }
  g_scratch_ptr[device] = (char*)p;
}

char* scratch_region(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end()) fail(SVM_ERR_STATE, "scratch region of device " + std::to_string(device) + " not set up");
  return it->second;
}

}  // namespace svm
