// groupsort.hip — group-by over key spaces a dense group table cannot hold.
//
// The reference reduces each bulk's passing rows with thrust::sort_by_key on the
// zipped key columns followed by thrust::reduce_by_key (core/pipeline.cu:380-471),
// so any int32 keys and any number of groups work.  The scan's dense mode indexes a
// table by the mixed-radix group id, which caps the group space; beyond it the scan
// packs each passing row's key codes into one 64-bit key (first key in the high
// bits, so the packed order is the lexicographic key order), pre-aggregates in an
// LDS hash table per workgroup and appends (packed key, sums) entries to HBM when a
// row finds no LDS slot and when a workgroup flushes.  Here those entries are sorted
// by key (rocPRIM radix sort), equal keys reduced (rocPRIM reduce_by_key, int64 sums)
// and the keys unpacked back into int32 columns.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>

#include "kernels.h"
#include "svm_internal.h"

namespace svm {
namespace k {

namespace {

constexpr int kBlock = 256;

unsigned grid_for(uint64_t n, unsigned cap = 65536) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g == 0) g = 1;
  return (unsigned)(g > cap ? cap : g);
}

__global__ void k_iota_u32(uint32_t* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

// out[i] = vals[perm[i] * M + m]
__global__ void k_gather_metric(const int64_t* __restrict__ vals, const uint32_t* __restrict__ perm, uint64_t n,
                                int M, int m, int64_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = vals[(uint64_t)perm[i] * (uint64_t)M + (uint64_t)m];
}

__global__ void k_unpack_keys(const uint64_t* __restrict__ keys, uint64_t G, UnpackArgs a) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[i];
    for (int q = 0; q < a.K; ++q) {
      const uint64_t mask = a.bits[q] >= 64 ? ~0ull : ((1ull << a.bits[q]) - 1ull);
      const uint32_t code = (uint32_t)((key >> a.shift[q]) & mask);
      // a dimension key's code indexes its sorted dictionary; a source-column key's
      // code is its value minus the column minimum (int32 wrap-around)
      a.out[q][i] = a.dict[q] ? a.dict[q][code] : (int32_t)(code + (uint32_t)a.fmin[q]);
    }
  }
}

}  // namespace

uint64_t sort_reduce(const uint64_t* keys, const int64_t* vals, uint64_t n, int M, int key_bits,
                     uint64_t* uniq_keys, int64_t* const* sums, hipStream_t s) {
  if (!n) return 0;
  if (M < 1 || M > 8) fail(SVM_ERR_INVALID, "sort_reduce: 1..8 metrics");
  if (n > 0xFFFFFFFFull) fail(SVM_ERR_UNSUPPORTED, "more than 2^32 group entries to sort");
  const int end_bit = key_bits < 1 ? 1 : key_bits > 64 ? 64 : key_bits;
  DevBuf perm_in(n * 4 + 4, s), perm(n * 4 + 4, s), skeys(n * 8 + 8, s), sval(n * 8 + 8, s), cnt(16, s);
  k_iota_u32<<<grid_for(n), kBlock, 0, s>>>(perm_in.as<uint32_t>(), n);
  hip_check(hipGetLastError(), "sort_reduce iota");
  size_t t_sort = 0, t_red = 0;
  SVM_HIP(rocprim::radix_sort_pairs(nullptr, t_sort, keys, skeys.as<uint64_t>(), perm_in.as<uint32_t>(),
                                    perm.as<uint32_t>(), (size_t)n, 0, end_bit, s));
  SVM_HIP(rocprim::reduce_by_key(nullptr, t_red, skeys.as<uint64_t>(), sval.as<int64_t>(), (size_t)n, uniq_keys,
                                 sums[0], cnt.as<unsigned long long>(), rocprim::plus<int64_t>(),
                                 rocprim::equal_to<uint64_t>(), s));
  DevBuf tmp((t_sort > t_red ? t_sort : t_red) + 256, s);
  SVM_HIP(rocprim::radix_sort_pairs(tmp.p, t_sort, keys, skeys.as<uint64_t>(), perm_in.as<uint32_t>(),
                                    perm.as<uint32_t>(), (size_t)n, 0, end_bit, s));
  for (int m = 0; m < M; ++m) {
    k_gather_metric<<<grid_for(n), kBlock, 0, s>>>(vals, perm.as<uint32_t>(), n, M, m, sval.as<int64_t>());
    hip_check(hipGetLastError(), "sort_reduce gather");
    SVM_HIP(rocprim::reduce_by_key(tmp.p, t_red, skeys.as<uint64_t>(), sval.as<int64_t>(), (size_t)n, uniq_keys,
                                   sums[m], cnt.as<unsigned long long>(), rocprim::plus<int64_t>(),
                                   rocprim::equal_to<uint64_t>(), s));
  }
  unsigned long long h = 0;
  SVM_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  return h;
}

void unpack_keys(const uint64_t* keys, uint64_t G, const UnpackArgs& a, hipStream_t s) {
  if (!G) return;
  k_unpack_keys<<<grid_for(G), kBlock, 0, s>>>(keys, G, a);
  hip_check(hipGetLastError(), "unpack_keys");
}

}  // namespace k
}  // namespace svm
