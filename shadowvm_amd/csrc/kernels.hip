// kernels.hip — precompiled gfx950 kernels: dense-map build, dictionaries, group
// emission, scans and the synthetic table generator.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "kernels.h"
#include "svm_internal.h"

namespace svm {
namespace k {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n, int per_block = kBlock, unsigned cap = 65536) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g == 0) g = 1;
  return (unsigned)(g > cap ? cap : g);
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_minmax_init(int32_t* out) {
  out[0] = INT_MAX;
  out[1] = INT_MIN;
}

__global__ void k_minmax(const int32_t* __restrict__ col, uint64_t n, int32_t* out) {
  int lo = INT_MAX, hi = INT_MIN;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    int v = col[i];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int d = 32; d > 0; d >>= 1) {
    int a = __shfl_xor(lo, d), b = __shfl_xor(hi, d);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&out[0], lo);
    atomicMax(&out[1], hi);
  }
}

// Narrow copy of an Int32 column, in blocks of 512 rows: the low 16 bits of the
// block's values v - kmin (< 2^(16 + HB)) as 512 uint16, then their HB high bits
// (HB = 0, 2 or 8) packed row by row (1,024 + 64 x HB bytes per block).  A wave of a
// scan then reads 512 rows as one coalesced 1 KB load and one coalesced load of 64 x
// HB bytes.  One thread per 8 rows: two 128-bit loads, one 128-bit store and one store
// of its 8 x HB high bits.
template <int HB>
__global__ void k_pack_keys(const int32_t* __restrict__ col, uint64_t n, int32_t kmin, uint8_t* __restrict__ out) {
  constexpr uint64_t kBlockBytes = 1024 + 64 * HB;
  constexpr unsigned kMask = (1u << (16 + HB)) - 1u;
  const uint64_t groups = (n + 7) / 8;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += (uint64_t)gridDim.x * blockDim.x) {
    const int4* src = (const int4*)col + 2 * g;  // the column padding covers a partial last group
    const int4 a = src[0], b = src[1];
    const int v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    unsigned x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = 8 * g + (uint64_t)j < n ? (unsigned)(v[j] - kmin) & kMask : 0u;
    uint8_t* blk = out + kBlockBytes * (g >> 6);
    *(uint4*)(blk + 16 * (g & 63)) = make_uint4((x[0] & 0xFFFFu) | (x[1] << 16), (x[2] & 0xFFFFu) | (x[3] << 16),
                                                (x[4] & 0xFFFFu) | (x[5] << 16), (x[6] & 0xFFFFu) | (x[7] << 16));
    if (HB == 8) {
      *(uint2*)(blk + 1024 + 8 * (g & 63)) =
          make_uint2((x[0] >> 16) | (x[1] >> 16 << 8) | (x[2] >> 16 << 16) | (x[3] >> 16 << 24),
                     (x[4] >> 16) | (x[5] >> 16 << 8) | (x[6] >> 16 << 16) | (x[7] >> 16 << 24));
    } else if (HB == 2) {
      unsigned h = 0u;
#pragma unroll
      for (int j = 0; j < 8; ++j) h |= (x[j] >> 16) << (2 * j);
      *(uint16_t*)(blk + 1024 + 2 * (g & 63)) = (uint16_t)h;
    }
  }
}

__device__ inline unsigned hash31(const char* p, int n) {
  unsigned h = 0;
  for (int q = 0; q < n; ++q) h = h * 31u + (unsigned)(int)(signed char)p[q];
  return h & 0x7FFFFFFFu;
}

__global__ void k_values(const char* __restrict__ col, int type, int64_t size, uint64_t n,
                         int32_t* vals) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const char* p = col + r * (uint64_t)size;
    int32_t v;
    switch (type) {
      case Byte: v = (int32_t)*(const signed char*)p; break;
      case Int16: v = (int32_t)*(const short*)p; break;
      case String: v = (int32_t)hash31(p, (int)size); break;
      default: v = *(const int32_t*)p; break;
    }
    vals[r] = v;
  }
}

__global__ void k_codes(const int32_t* __restrict__ vals, uint64_t n, const int32_t* __restrict__ dict,
                        uint64_t nd, int32_t* codes) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    int32_t v = vals[r];
    uint64_t lo = 0, hi = nd;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (dict[mid] < v) lo = mid + 1; else hi = mid;
    }
    codes[r] = (int32_t)lo;
  }
}

__global__ void k_first_rows(const int32_t* __restrict__ keys, uint64_t n, int32_t kmin, uint32_t range,
                             int32_t* first, const uint8_t* __restrict__ kept) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t off = (uint32_t)(keys[r] - kmin);
    if (off < range && (!kept || kept[r])) atomicMin(&first[off], (int32_t)r);
  }
}

// *flag = 1 when two rows share a key.  One bit per key offset, set with atomicOr; a
// plain read first skips the atomic for keys already seen (a set bit can only come from
// another row, and a stale 0 falls through to the exact atomic), so a column with few
// distinct keys does not serialise on them.  Threads stop once the flag is up.
__global__ void k_dup_bits(const int32_t* __restrict__ keys, uint64_t n, int32_t kmin, uint32_t range,
                           uint32_t* bits, uint32_t* flag) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const uint32_t off = (uint32_t)(keys[r] - kmin);
    if (off >= range) continue;
    const uint32_t b = 1u << (off & 31u);
    uint32_t* w = bits + (off >> 5);
    if ((*w & b) || (atomicOr(w, b) & b)) atomicExch(flag, 1u);
  }
}

__global__ void k_pack(const int32_t* __restrict__ first, uint32_t range, int payload,
                       const int32_t* __restrict__ codes, int wbits, uint32_t* packed, uint64_t nwords) {
  const int per = 32 / wbits;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t word = 0;
    for (int e = 0; e < per; ++e) {
      uint64_t slot = w * (uint64_t)per + (uint64_t)e;
      if (slot >= range) break;
      int32_t f = first[slot];
      uint32_t val = 0;
      if (f != INT_MAX) val = payload == 0 ? 1u : payload == 1 ? (uint32_t)codes[f] + 1u : (uint32_t)f + 1u;
      word |= wbits == 32 ? val : (val << (e * wbits));
    }
    packed[w] = word;
  }
}

// The packed map of a dimension object whose keys are unique: every row is its key's
// first row, so each row writes its own slot (payload 0: existence, 1: code + 1 of
// the row's attribute value in the sorted dictionary, 2: row + 1).  The map must be
// zeroed; slots narrower than 32 bits share words and are set with atomicOr.
__global__ void k_pack_direct(const int32_t* __restrict__ keys, uint64_t n, int32_t kmin, uint32_t range,
                              int payload, const char* __restrict__ attr, int atype, int64_t asize,
                              const int32_t* __restrict__ dict, uint64_t nd, int wbits, uint32_t* packed,
                              uint32_t* bits) {
  const unsigned per = 32u / (unsigned)wbits;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t off = (uint32_t)(keys[r] - kmin);
    if (off >= range) continue;
    uint32_t val = 1u;
    if (payload == 2) {
      val = (uint32_t)r + 1u;
    } else if (payload == 1) {
      const char* p = attr + r * (uint64_t)asize;
      int32_t v;
      switch (atype) {
        case Byte: v = (int32_t)*(const signed char*)p; break;
        case Int16: v = (int32_t)*(const short*)p; break;
        case String: v = (int32_t)hash31(p, (int)asize); break;
        default: v = *(const int32_t*)p; break;
      }
      uint64_t lo = 0, hi = nd;
      while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (dict[mid] < v) lo = mid + 1; else hi = mid;
      }
      val = (uint32_t)lo + 1u;
    }
    if (wbits == 32) packed[off] = val;
    else atomicOr(&packed[off / per], val << ((off % per) * (unsigned)wbits));
    if (bits) atomicOr(&bits[off >> 5], 1u << (off & 31u));
  }
}

// Existence bits of a packed map: bit s of bits[] is set when slot s holds a payload.
__global__ void k_exist_bits(const uint32_t* __restrict__ packed, int wbits, uint32_t range, uint32_t* bits,
                             uint64_t nbw) {
  const unsigned per = 32u / (unsigned)wbits;
  const uint32_t mask = wbits == 32 ? 0xFFFFFFFFu : (1u << wbits) - 1u;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nbw;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t out = 0;
    for (unsigned e = 0; e < 32u; ++e) {
      const uint64_t slot = w * 32u + e;
      if (slot >= range) break;
      const uint32_t v = wbits == 32 ? packed[slot] : (packed[slot / per] >> ((slot % per) * (unsigned)wbits)) & mask;
      out |= (v != 0u ? 1u : 0u) << e;
    }
    bits[w] = out;
  }
}

// Single-workgroup exclusive scan with a running carry (inputs up to a few million).
template <typename T>
__global__ void k_exscan(const T* __restrict__ in, uint64_t n, unsigned long long* out,
                         unsigned long long* total, unsigned long long* host_total) {
  __shared__ unsigned long long buf[1024];
  unsigned long long carry = 0;
  for (uint64_t base = 0; base < n; base += 1024) {
    uint64_t i = base + threadIdx.x;
    unsigned long long v = i < n ? (unsigned long long)in[i] : 0ull;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      unsigned long long t = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0ull;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) out[i] = carry + buf[threadIdx.x] - v;
    unsigned long long blk = buf[1023];
    __syncthreads();
    carry += blk;
  }
  if (threadIdx.x == 0) {
    *total = carry;
    if (host_total) *host_total = carry;
  }
}

__global__ void k_group_popc(const uint32_t* __restrict__ flags, uint64_t nwords,
                             unsigned long long* counts) {
  __shared__ unsigned part[kBlock / 64];
  uint64_t w = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned c = w < nwords ? (unsigned)__popc(flags[w]) : 0u;
  for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int q = 0; q < kBlock / 64; ++q) t += part[q];
    counts[blockIdx.x] = t;
  }
}

// One thread per group id: block b covers groups [256 b, 256 b + 256), i.e. flag
// words [8 b, 8 b + 8).  `offsets` holds the exclusive prefix of set flags per
// k_group_popc block (kBlock words); the rows of earlier words inside that block
// are counted here, then a wave ballot orders the block's own groups.
__global__ void k_group_emit(EmitArgs a, const unsigned long long* __restrict__ offsets) {
  static_assert(kBlock == 256, "one block covers 8 flag words");
  __shared__ unsigned red[kBlock / 64];
  __shared__ unsigned cnt[kBlock / 64];
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t w0 = (uint64_t)blockIdx.x * (kBlock / 32);
  const uint64_t wa = w0 / kBlock * kBlock;  // first word of the k_group_popc block
  unsigned c = wa + threadIdx.x < w0 ? (unsigned)__popc(a.flags[wa + threadIdx.x]) : 0u;
  for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
  if (lane == 0) red[wave] = c;
  const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool set = g < a.G && ((a.flags[g >> 5] >> (g & 31)) & 1u);
  const unsigned long long m = __ballot(set);
  if (lane == 0) cnt[wave] = (unsigned)__popcll(m);
  // the padding behind the last row of every column (rows are written below it only)
  if (a.total && blockIdx.x == 0 && threadIdx.x < kPadBytes) {
    const unsigned long long n = *a.total;
    for (int q = 0; q < a.K; ++q) ((char*)(a.keys_out[q] + n))[threadIdx.x] = 0;
    for (int m2 = 0; m2 < a.M; ++m2) ((char*)(a.sums_out[m2] + n))[threadIdx.x] = 0;
  }
  __syncthreads();
  if (!set) return;
  unsigned long long row = offsets[w0 / kBlock] + red[0] + red[1] + red[2] + red[3];
  for (unsigned q = 0; q < wave; ++q) row += cnt[q];
  row += (unsigned)__popcll(m & ((1ull << lane) - 1ull));
  for (int q = 0; q < a.K; ++q) {
    uint32_t code = (uint32_t)((g / a.stride[q]) % a.n[q]);
    const int32_t v = a.dict[q] ? a.dict[q][code] : a.fmin[q] + (int32_t)code;
    a.keys_out[q][row] = v;
    if (a.keys_host[q]) a.keys_host[q][row] = v;
  }
  for (int m2 = 0; m2 < a.M; ++m2) {
    const int64_t v = (int64_t)a.sums[g * (uint64_t)a.M + m2];
    a.sums_out[m2][row] = v;
    if (a.sums_host[m2]) a.sums_host[m2][row] = v;
  }
}

struct ZeroRanges {
  uint32_t* p[4];
  uint64_t n[4];
};
__global__ void k_zero_ranges(ZeroRanges z) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int r = 0; r < 4; ++r)
    for (uint64_t i = t; i < z.n[r]; i += stride) z.p[r][i] = 0u;
}

// *flag = 1 when a[i] != b[i] for some i < n.
__global__ void k_words_differ(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t n,
                               uint32_t* flag) {
  uint32_t d = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    d |= a[i] ^ b[i];
  if (__any(d != 0u) && (threadIdx.x & 63u) == 0u) atomicOr(flag, 1u);
}

__global__ void k_add_u64(unsigned long long* __restrict__ dst, const unsigned long long* __restrict__ src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] += src[i];
}
__global__ void k_or_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] |= src[i];
}

// Zeroes kColumnPad bytes at each pointer (one block per pointer).
__global__ void k_zero_pads(PadList l) {
  if (blockIdx.x < (unsigned)l.n && threadIdx.x < kPadBytes) l.p[blockIdx.x][threadIdx.x] = 0;
}

__global__ void svm_calib_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Streams `n16` 16-byte groups of each of `ncols` columns with non-temporal 128-bit
// loads (the access pattern of the scans) and nothing else: what HBM gives a read of
// those columns at a launch shape.  The XOR of everything read goes to out[0] only in
// a case that cannot happen for the data here, so no load is optimised away.
typedef int i32x4 __attribute__((ext_vector_type(4)));
struct ReadCols {
  const i32x4* c[8];
};
__global__ void __launch_bounds__(1024) svm_read_probe(ReadCols cols, int ncols, uint64_t n16, i32x4* out) {
  i32x4 acc = {0, 0, 0, 0};
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n16; v += stride) {
    for (int k = 0; k < 8; ++k) {
      if (k >= ncols) break;
      acc ^= __builtin_nontemporal_load(cols.c[k] + v);
    }
  }
  if (acc.x == 0x7A3B9C1D && acc.y == 0x5E2F8A47 && acc.z == 0x1C6D3E9B && acc.w == 0x4F8B2A6C) out[0] = acc;
}

struct GenCols {
  void* c[SVM_GEN_MAX_COLS];
};

__global__ void k_gen(svm_gen_ctx ctx, int kind, uint64_t row0, uint64_t n, GenCols cols) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x)
    svm_gen_row(&ctx, kind, row0 + r, r, cols.c);
}

}  // namespace

#define LAUNCH_CHECK(what) ::svm::hip_check(hipGetLastError(), what)

void fill_u32(void* p, uint64_t nwords, uint32_t v, hipStream_t s) {
  if (!nwords) return;
  k_fill_u32<<<grid_for(nwords), kBlock, 0, s>>>((uint32_t*)p, nwords, v);
  LAUNCH_CHECK("fill_u32");
}

void pack_keys(const int32_t* col, uint64_t n, int32_t kmin, int hb, uint8_t* out, hipStream_t s) {
  if (!n) return;
  const auto grid = grid_for((n + 7) / 8, kBlock, 4096);
  if (hb == 8) k_pack_keys<8><<<grid, kBlock, 0, s>>>(col, n, kmin, out);
  else if (hb == 2) k_pack_keys<2><<<grid, kBlock, 0, s>>>(col, n, kmin, out);
  else k_pack_keys<0><<<grid, kBlock, 0, s>>>(col, n, kmin, out);
  LAUNCH_CHECK("pack_keys");
}

void minmax_i32(const int32_t* col, uint64_t n, int32_t* out, hipStream_t s) {
  k_minmax_init<<<1, 1, 0, s>>>(out);
  if (n) k_minmax<<<grid_for(n, kBlock, 2048), kBlock, 0, s>>>(col, n, out);
  LAUNCH_CHECK("minmax_i32");
}

void values_of(const void* col, int type, int64_t size, uint64_t n, int32_t* vals, hipStream_t s) {
  if (!n) return;
  k_values<<<grid_for(n), kBlock, 0, s>>>((const char*)col, type, size, n, vals);
  LAUNCH_CHECK("values_of");
}

uint64_t dict_build(const int32_t* vals, uint64_t n, int32_t* dict, hipStream_t s) {
  if (!n) return 0;
  DevBuf sorted(n * sizeof(int32_t));
  size_t t1 = 0, t2 = 0;
  SVM_HIP(rocprim::radix_sort_keys(nullptr, t1, vals, sorted.as<int32_t>(), n, 0, 32, s));
  SVM_HIP(rocprim::unique(nullptr, t2, sorted.as<int32_t>(), dict, (unsigned long long*)nullptr, n,
                          rocprim::equal_to<int32_t>(), s));
  DevBuf tmp((t1 > t2 ? t1 : t2) + 256);
  DevBuf cnt(sizeof(unsigned long long));
  SVM_HIP(rocprim::radix_sort_keys(tmp.p, t1, vals, sorted.as<int32_t>(), n, 0, 32, s));
  SVM_HIP(rocprim::unique(tmp.p, t2, sorted.as<int32_t>(), dict, cnt.as<unsigned long long>(), n,
                          rocprim::equal_to<int32_t>(), s));
  unsigned long long h = 0;
  SVM_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  return h;
}

void codes_of(const int32_t* vals, uint64_t n, const int32_t* dict, uint64_t nd, int32_t* codes,
              hipStream_t s) {
  if (!n) return;
  k_codes<<<grid_for(n), kBlock, 0, s>>>(vals, n, dict, nd, codes);
  LAUNCH_CHECK("codes_of");
}

void first_rows(const int32_t* keys, uint64_t n, int32_t kmin, uint32_t range, int32_t* first,
                hipStream_t s, const uint8_t* kept) {
  fill_u32(first, range, (uint32_t)INT_MAX, s);
  if (n) k_first_rows<<<grid_for(n), kBlock, 0, s>>>(keys, n, kmin, range, first, kept);
  LAUNCH_CHECK("first_rows");
}

void exist_bits(const uint32_t* packed, int wbits, uint32_t range, uint32_t* bits, hipStream_t s) {
  const uint64_t nbw = exist_words(range);
  if (!nbw) return;
  k_exist_bits<<<grid_for(nbw), kBlock, 0, s>>>(packed, wbits, range, bits, nbw);
  LAUNCH_CHECK("exist_bits");
}

void pack_direct(const int32_t* keys, uint64_t n, int32_t kmin, uint32_t range, int payload, const void* attr,
                 int atype, int64_t asize, const int32_t* dict, uint64_t nd, int wbits, uint32_t* packed,
                 uint32_t* bits, hipStream_t s) {
  const uint64_t per = 32 / (uint64_t)wbits;
  const uint64_t words = ((uint64_t)range + per - 1) / per;
  // the existence bits, when kept, follow the packed words: both cleared at once
  fill_u32(packed, words + (bits ? exist_words(range) : 0), 0u, s);
  if (n) k_pack_direct<<<grid_for(n), kBlock, 0, s>>>(keys, n, kmin, range, payload, (const char*)attr, atype,
                                                     asize, dict, nd, wbits, packed, bits);
  LAUNCH_CHECK("pack_direct");
}

void pack_map(const int32_t* first, uint32_t range, int payload, const int32_t* codes, int wbits,
              uint32_t* packed, hipStream_t s) {
  uint64_t per = 32 / (uint64_t)wbits;
  uint64_t nwords = ((uint64_t)range + per - 1) / per;
  if (!nwords) return;
  k_pack<<<grid_for(nwords), kBlock, 0, s>>>(first, range, payload, codes, wbits, packed, nwords);
  LAUNCH_CHECK("pack_map");
}

bool has_duplicates(const int32_t* keys, uint64_t n, int32_t kmin, uint32_t range, hipStream_t s) {
  if (!n || !range) return false;
  const uint64_t nbw = ((uint64_t)range + 31) / 32;
  DevBuf bits(nbw * 4 + 4);
  uint32_t* flag = bits.as<uint32_t>() + nbw;
  fill_u32(bits.p, nbw + 1, 0, s);
  k_dup_bits<<<grid_for(n), kBlock, 0, s>>>(keys, n, kmin, range, bits.as<uint32_t>(), flag);
  LAUNCH_CHECK("has_duplicates");
  uint32_t h = 0;
  SVM_HIP(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  return h != 0;
}

void exclusive_scan_u32(const uint32_t* in, uint64_t n, unsigned long long* out,
                        unsigned long long* total, hipStream_t s, unsigned long long* host_total) {
  k_exscan<uint32_t><<<1, 1024, 0, s>>>(in, n, out, total, host_total);
  LAUNCH_CHECK("exclusive_scan_u32");
}

uint64_t group_count(const uint32_t* flags, uint64_t G, unsigned long long* blocks,
                     unsigned long long* offsets, hipStream_t s) {
  uint64_t nwords = (G + 31) / 32;
  uint64_t nb = (nwords + kBlock - 1) / kBlock;
  if (!nb) return 0;
  k_group_popc<<<(unsigned)nb, kBlock, 0, s>>>(flags, nwords, blocks);
  const HostSlots hs = host_slots(s);
  k_exscan<unsigned long long><<<1, 1024, 0, s>>>(blocks, nb, offsets, offsets + nb, hs.dev + 2);
  LAUNCH_CHECK("group_count");
  SVM_HIP(hipStreamSynchronize(s));
  return hs.host[2];
}

uint64_t group_count_async(const uint32_t* flags, uint64_t G, unsigned long long* blocks,
                           unsigned long long* offsets, unsigned long long* count_out, hipStream_t s) {
  const uint64_t nwords = (G + 31) / 32;
  const uint64_t nb = (nwords + kBlock - 1) / kBlock;
  if (!nb) fail(SVM_ERR_INVALID, "group_count_async: no groups");
  k_group_popc<<<(unsigned)nb, kBlock, 0, s>>>(flags, nwords, blocks);
  k_exscan<unsigned long long><<<1, 1024, 0, s>>>(blocks, nb, offsets, offsets + nb, count_out);
  LAUNCH_CHECK("group_count_async");
  return nb;
}

void group_emit(const EmitArgs& a, const unsigned long long* offsets, hipStream_t s) {
  const uint64_t nb = (a.G + kBlock - 1) / kBlock;
  if (!nb) return;
  if (nb > 0x7FFFFFFFull) fail(SVM_ERR_UNSUPPORTED, "group_emit: group space too large");
  k_group_emit<<<(unsigned)nb, kBlock, 0, s>>>(a, offsets);
  LAUNCH_CHECK("group_emit");
}

void zero_ranges(void* const* ptrs, const uint64_t* words, int n, hipStream_t s) {
  if (n < 0 || n > 4) fail(SVM_ERR_INVALID, "zero_ranges: at most 4 ranges");
  ZeroRanges z;
  uint64_t most = 0;
  for (int r = 0; r < 4; ++r) {
    z.p[r] = r < n ? (uint32_t*)ptrs[r] : nullptr;
    z.n[r] = r < n && ptrs[r] ? words[r] : 0;
    most = z.n[r] > most ? z.n[r] : most;
  }
  if (!most) return;
  k_zero_ranges<<<grid_for(most), kBlock, 0, s>>>(z);
  LAUNCH_CHECK("zero_ranges");
}

void add_u64(void* dst, const void* src, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_add_u64<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>((unsigned long long*)dst, (const unsigned long long*)src, n);
  LAUNCH_CHECK("add_u64");
}

void or_u32(void* dst, const void* src, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_or_u32<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>((uint32_t*)dst, (const uint32_t*)src, n);
  LAUNCH_CHECK("or_u32");
}

bool words_differ(const void* a, const void* b, uint64_t nwords, uint32_t* flag, hipStream_t s) {
  if (!nwords) return false;
  k_fill_u32<<<1, 64, 0, s>>>(flag, 1, 0u);
  k_words_differ<<<grid_for(nwords, kBlock, 4096), kBlock, 0, s>>>((const uint32_t*)a, (const uint32_t*)b, nwords,
                                                                    flag);
  LAUNCH_CHECK("words_differ");
  uint32_t h = 0;
  SVM_HIP(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  return h != 0;
}

void zero_pads(char* const* ptrs, int n, hipStream_t s) {
  for (int i = 0; i < n; i += kMaxPads) {
    PadList l;
    l.n = n - i < kMaxPads ? n - i : kMaxPads;
    for (int q = 0; q < l.n; ++q) l.p[q] = ptrs[i + q];
    k_zero_pads<<<(unsigned)l.n, kPadBytes, 0, s>>>(l);
    LAUNCH_CHECK("zero_pads");
  }
}

void read_probe(const void* const* cols, int ncols, uint64_t rows, int block, unsigned grid, void* out,
                hipStream_t s) {
  if (ncols < 1 || ncols > 8) fail(SVM_ERR_INVALID, "read_probe: 1..8 columns");
  if (block < 64 || block > 1024 || block % 64) fail(SVM_ERR_INVALID, "read_probe: block of 64..1024 threads");
  ReadCols c;
  memset(&c, 0, sizeof(c));
  for (int k = 0; k < ncols; ++k) c.c[k] = (const i32x4*)cols[k];
  const uint64_t n16 = rows / 4;
  if (!n16) return;
  svm_read_probe<<<grid ? grid : 1, block, 0, s>>>(c, ncols, n16, (i32x4*)out);
  LAUNCH_CHECK("read_probe");
}

void calib_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s) {
  uint64_t n16 = bytes / 16;
  if (!n16) return;
  svm_calib_copy<<<grid_for(n16, kBlock, 8192), kBlock, 0, s>>>((const uint4*)src, (uint4*)dst, n16);
  LAUNCH_CHECK("calib_copy");
}

void gen_table(const svm_gen_ctx& ctx, int kind, uint64_t row0, uint64_t n, void* const* cols,
               hipStream_t s) {
  if (!n) return;
  GenCols gc;
  for (int c = 0; c < SVM_GEN_MAX_COLS; ++c) gc.c[c] = cols[c];
  k_gen<<<grid_for(n, kBlock, 16384), kBlock, 0, s>>>(ctx, kind, row0, n, gc);
  LAUNCH_CHECK("gen_table");
}

}  // namespace k
}  // namespace svm
