// loader_kernels.hip — parsing '|'-separated text in HBM into columns.
//
// Rows are found with a chunked count + scan + scatter of the row separator's
// positions; then one thread parses one row: fields split at the column
// separator, integers as the C library's strtol does (leading white space,
// optional sign, decimal digits, saturation on overflow; atoi truncates that to
// int), strings copied and NUL-terminated at their declared width
// (com/cpp/com.cu:33-44 GetValueOfItemBySwitch).
#include <hip/hip_runtime.h>

#include <climits>

#include "loader.h"
#include "svm_internal.h"

namespace svm {
namespace ld {

namespace {

constexpr int kBlock = 256;
constexpr int kPerThread = (int)(kChunk / kBlock);  // 64 bytes = 4 x 16
static_assert(kPerThread == 64, "chunk geometry");

// Separators in the 64 bytes of this thread's segment (bytes at or past n do not count).
__device__ __forceinline__ unsigned segment_mask(const char* buf, uint64_t n, char sep, uint64_t s,
                                                 unsigned long long& bits) {
  bits = 0ull;
  if (s >= n) return 0u;
  const uint4* v = (const uint4*)(buf + s);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint4 w = v[j];
    const unsigned words[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const unsigned char ch = (unsigned char)(words[k >> 2] >> (8 * (k & 3)));
      const uint64_t p = s + (uint64_t)(j * 16 + k);
      if (ch == (unsigned char)sep && p < n) bits |= 1ull << (j * 16 + k);
    }
  }
  return (unsigned)__popcll(bits);
}

__global__ void k_count_sep(const char* __restrict__ buf, uint64_t n, char sep, unsigned* __restrict__ counts,
                            uint64_t nchunks) {
  __shared__ unsigned part[kBlock / 64];
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    unsigned long long bits;
    unsigned cnt = segment_mask(buf, n, sep, c * kChunk + (uint64_t)threadIdx.x * kPerThread, bits);
    for (int d = 32; d > 0; d >>= 1) cnt += __shfl_xor(cnt, d);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned t = 0;
      for (int w = 0; w < kBlock / 64; ++w) t += part[w];
      counts[c] = t;
    }
    __syncthreads();
  }
}

__global__ void k_sep_pos(const char* __restrict__ buf, uint64_t n, char sep,
                          const unsigned long long* __restrict__ offsets, unsigned long long* __restrict__ ends,
                          uint64_t nchunks) {
  __shared__ unsigned pre[kBlock];
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const uint64_t s = c * kChunk + (uint64_t)threadIdx.x * kPerThread;
    unsigned long long bits;
    const unsigned cnt = segment_mask(buf, n, sep, s, bits);
    pre[threadIdx.x] = cnt;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
      const unsigned t = threadIdx.x >= (unsigned)d ? pre[threadIdx.x - d] : 0u;
      __syncthreads();
      pre[threadIdx.x] += t;
      __syncthreads();
    }
    unsigned long long o = offsets[c] + pre[threadIdx.x] - cnt;
    while (bits) {
      const int b = __ffsll((long long)bits) - 1;
      bits &= bits - 1ull;
      ends[o++] = s + (uint64_t)b;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ bool c_space(char ch) {
  return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\v' || ch == '\f' || ch == '\r';
}

// strtol(item, NULL, 10) over the field bytes [p, e) (the item is NUL-terminated
// at the separator, so a NUL byte ends the digits like any other non-digit).
__device__ long long parse_long(const char* p, const char* e) {
  while (p < e && c_space(*p)) ++p;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    ++p;
  }
  const unsigned long long lim = neg ? 9223372036854775808ull : 9223372036854775807ull;
  unsigned long long acc = 0;
  bool over = false, any = false;
  for (; p < e && *p >= '0' && *p <= '9'; ++p) {
    const unsigned d = (unsigned)(*p - '0');
    any = true;
    if (over || acc > (lim - d) / 10ull) over = true;
    else acc = acc * 10ull + d;
  }
  if (!any) return 0;
  if (over) return neg ? LLONG_MIN : LLONG_MAX;
  return neg ? (long long)(0ull - acc) : (long long)acc;
}

__device__ void store_field(const ParseArgs& a, int col, uint64_t row, const char* p, const char* e) {
  char* dst = (char*)a.col[col] + row * (uint64_t)a.size[col];
  switch (a.type[col]) {
    case Byte: *dst = p < e ? *p : '\0'; break;
    case Int16: *(short*)dst = (short)(int)parse_long(p, e); break;
    case Int32: *(int*)dst = (int)parse_long(p, e); break;
    case Int64: *(long long*)dst = parse_long(p, e); break;
    default: {  // String: the item, NUL-padded, last byte NUL
      const int size = a.size[col];
      const long long len = e - p;
      for (int k = 0; k < size - 1; ++k) dst[k] = k < len ? p[k] : '\0';
      dst[size - 1] = '\0';
    }
  }
}

__global__ void k_parse(ParseArgs a) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.rows;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t start = r == 0 ? 0ull : a.ends[r - 1] + 1ull;
    const uint64_t end = a.ends[r];
    const uint64_t row = a.row0 + r;
    int col = 0;
    uint64_t f = start;
    for (uint64_t p = start; p < end; ++p) {
      if (a.buf[p] == a.col_sep) {
        if (col < a.ncols) store_field(a, col, row, a.buf + f, a.buf + p);
        col++;
        f = p + 1;
      }
    }
    // the bytes after the last column separator are dropped at the row separator
  }
}

inline unsigned grid_for(uint64_t n, uint64_t per, unsigned cap) {
  uint64_t g = (n + per - 1) / per;
  if (g == 0) g = 1;
  return (unsigned)(g > cap ? cap : g);
}

}  // namespace

void count_separators(const char* buf, uint64_t n, char sep, unsigned* counts, uint64_t nchunks, hipStream_t s) {
  if (!nchunks) return;
  hipLaunchKernelGGL(k_count_sep, dim3(grid_for(nchunks, 1, 8192)), dim3(kBlock), 0, s, buf, n, sep, counts,
                     nchunks);
  hip_check(hipGetLastError(), "k_count_sep");
}

void separator_positions(const char* buf, uint64_t n, char sep, const unsigned long long* offsets,
                         unsigned long long* ends, uint64_t nchunks, hipStream_t s) {
  if (!nchunks) return;
  hipLaunchKernelGGL(k_sep_pos, dim3(grid_for(nchunks, 1, 8192)), dim3(kBlock), 0, s, buf, n, sep, offsets, ends,
                     nchunks);
  hip_check(hipGetLastError(), "k_sep_pos");
}

void parse_rows(const ParseArgs& a, hipStream_t s) {
  if (!a.rows) return;
  hipLaunchKernelGGL(k_parse, dim3(grid_for(a.rows, kBlock, 65536)), dim3(kBlock), 0, s, a);
  hip_check(hipGetLastError(), "k_parse");
}

}  // namespace ld
}  // namespace svm
