// kernels.h — precompiled gfx950 kernels of libsvm (build, finish and data generation
// steps that do not depend on the plan's expressions).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "datagen.h"

namespace svm {
namespace k {

// Every kernel below runs on `s` and returns without synchronising unless noted.
void fill_u32(void* p, uint64_t nwords, uint32_t v, hipStream_t s);
// out[0] = min, out[1] = max of col[0..n) (out[0..1] are initialised here).
void minmax_i32(const int32_t* col, uint64_t n, int32_t* out, hipStream_t s);
// Narrow copy of col (values in [kmin, kmin + 2^(16 + hb)), hb = 0, 2 or 8) in blocks
// of 512 rows (1,024 + 64 x hb bytes each): row i's value - kmin has its low 16 bits at
// uint16 i % 512 and its hb high bits at bit (i % 512) x hb of byte 1024 of block i / 512.
// Writes every row's bits; the unused rows of a partial last block are left as they were.
void pack_keys(const int32_t* col, uint64_t n, int32_t kmin, int hb, uint8_t* out, hipStream_t s);
// Group-key value of each dimension row: the int value, or Hash(bytes, size) for strings.
void values_of(const void* col, int type, int64_t size, uint64_t n, int32_t* vals, hipStream_t s);
// Sorted distinct values of vals[0..n) into dict; returns the count (synchronises).
uint64_t dict_build(const int32_t* vals, uint64_t n, int32_t* dict, hipStream_t s);
// codes[r] = position of vals[r] in dict (binary search).
void codes_of(const int32_t* vals, uint64_t n, const int32_t* dict, uint64_t nd, int32_t* codes,
              hipStream_t s);
// first[key-kmin] = smallest row r with keys[r] == key (first initialised to INT32_MAX),
// over the rows with kept[r] != 0 when kept is given.
void first_rows(const int32_t* keys, uint64_t n, int32_t kmin, uint32_t range, int32_t* first,
                hipStream_t s, const uint8_t* kept = nullptr);
// The reference's join hash table over keys[0..n) inserted in row order into `slots`
// slots (Hash of the key bytes % slots, at most 10 linear probes, silent drop):
// kept[r] = 1 if row r's insert found a slot, else 0.  Returns the number of dropped
// inserts (synchronises).
uint64_t ht_kept(const int32_t* keys, uint64_t n, uint64_t slots, uint8_t* kept, hipStream_t s);
// Packed map: slot = 0 if first == INT32_MAX else 1 / codes[row]+1 / row+1 by payload.
// bits[0 .. ceil(range/32)): bit s set when slot s of the packed map (wbits per slot)
// holds a payload.
void exist_bits(const uint32_t* packed, int wbits, uint32_t range, uint32_t* bits, hipStream_t s);
// Words of a dense map's existence bits over `range` slots: one bit per slot plus bit
// `range`, always 0, so a lookup may clamp its slot to `range` instead of testing it.
inline uint64_t exist_words(uint32_t range) { return (uint64_t)range / 32 + 1; }

// The packed map of a dimension object with unique keys, built in one pass over its
// rows (zeroes the map first).  payload 0: existence, 1: code + 1 of the row's value
// of attribute column `attr` (type atype, width asize; strings as Hash31) in the
// sorted dictionary dict[0..nd), 2: row + 1.
// bits (optional, = packed + its word count): the map's existence bits, set as well.
void pack_direct(const int32_t* keys, uint64_t n, int32_t kmin, uint32_t range, int payload, const void* attr,
                 int atype, int64_t asize, const int32_t* dict, uint64_t nd, int wbits, uint32_t* packed,
                 uint32_t* bits, hipStream_t s);
void pack_map(const int32_t* first, uint32_t range, int payload, const int32_t* codes, int wbits,
              uint32_t* packed, hipStream_t s);
// True when two rows of keys[0..n) share a key (keys within [kmin, kmin+range)); synchronises.
bool has_duplicates(const int32_t* keys, uint64_t n, int32_t kmin, uint32_t range, hipStream_t s);
// out[i] = sum(in[0..i)), *total = sum(in) (single workgroup); also writes the
// total to host_total (a pinned host word) when given.
void exclusive_scan_u32(const uint32_t* in, uint64_t n, unsigned long long* out,
                        unsigned long long* total, hipStream_t s, unsigned long long* host_total = nullptr);

struct EmitArgs {
  const uint32_t* flags;      // bit per group
  const uint64_t* sums;       // G x M
  uint64_t G;
  int K, M;
  uint32_t n[3], stride[3];   // cardinality and stride of each key component
  int32_t fmin[3];            // value of code 0 when dict is null
  const int32_t* dict[3];     // sorted dictionary of the component, or null
  int32_t* keys_out[3];       // K int32 columns
  int64_t* sums_out[8];       // M int64 columns
  // optional: the same rows also written here (host-mapped memory), and the row count
  // on the device, after which block 0 zeroes kColumnPad bytes of every output column
  int32_t* keys_host[3];
  int64_t* sums_host[8];
  const unsigned long long* total;
};
// Writes the groups whose flag is set, in group-id order, and returns their count
// (synchronises once to size the output).
uint64_t group_count(const uint32_t* flags, uint64_t G, unsigned long long* scratch_blocks,
                     unsigned long long* scratch_offsets, hipStream_t s);
// The same without synchronising: the count goes to *count_out (a device address, e.g.
// of a host-mapped word) and to scratch_offsets[blocks], blocks = (ceil(G/32)+255)/256;
// returns that offset.
uint64_t group_count_async(const uint32_t* flags, uint64_t G, unsigned long long* scratch_blocks,
                           unsigned long long* scratch_offsets, unsigned long long* count_out, hipStream_t s);
void group_emit(const EmitArgs& a, const unsigned long long* offsets, hipStream_t s);

// dst[i] += src[i] for n 64-bit words; dst[i] |= src[i] for n 32-bit words (merging
// a host engine's partial group table into the device's).
void add_u64(void* dst, const void* src, uint64_t n, hipStream_t s);
void or_u32(void* dst, const void* src, uint64_t n, hipStream_t s);

// True when the nwords 32-bit words at a and b differ (flag: one device word of
// scratch); synchronises.
bool words_differ(const void* a, const void* b, uint64_t nwords, uint32_t* flag, hipStream_t s);

// Group-by over wide key spaces (groupsort.hip).  keys[0..n) are packed group keys
// (key_bits significant bits), vals[i * M + m] their metric sums.  Sorts by key and
// sums equal keys: uniq_keys[0..G) ascending, sums[m][0..G); returns G (synchronises).
// Temporaries come from the stream's arena.
uint64_t sort_reduce(const uint64_t* keys, const int64_t* vals, uint64_t n, int M, int key_bits,
                     uint64_t* uniq_keys, int64_t* const* sums, hipStream_t s);
struct UnpackArgs {
  int K;
  int shift[3], bits[3];   // field of key q in the packed key
  const int32_t* dict[3];  // sorted dictionary of a dimension key, or null
  int32_t fmin[3];         // source-column key: value = code + fmin
  int32_t* out[3];         // K int32 key columns
};
void unpack_keys(const uint64_t* keys, uint64_t G, const UnpackArgs& a, hipStream_t s);

// Zeroes words[r] 32-bit words at ptrs[r] for r < n (n <= 4), in one launch.
void zero_ranges(void* const* ptrs, const uint64_t* words, int n, hipStream_t s);

// Zeroes the 256 padding bytes behind each of `n` column ends, in one launch per 32.
constexpr int kMaxPads = 32;
constexpr int kPadBytes = 256;
struct PadList {
  char* p[kMaxPads];
  int n;
};
void zero_pads(char* const* ptrs, int n, hipStream_t s);

// Non-temporal 128-bit reads of rows/4 groups of each of `ncols` (<= 8) int32 columns
// and nothing else (svm_read_probe), on `grid` workgroups of `block` threads; out: 16
// bytes of device scratch.
void read_probe(const void* const* cols, int ncols, uint64_t rows, int block, unsigned grid, void* out,
                hipStream_t s);
// Streaming 16-byte copy of `bytes` (a multiple of 16) — counter calibration.
void calib_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s);
// Synthetic table rows [row0, row0+n) of `kind` into device columns.
void gen_table(const svm_gen_ctx& ctx, int kind, uint64_t row0, uint64_t n, void* const* cols,
               hipStream_t s);

}  // namespace k
}  // namespace svm
