// refhash.hip — which rows the reference's join hash table keeps.
//
// The reference builds one table per join over bulk 0 of the dimension
// (core/pipeline.cu:88-133, com/cpp/com.hpp:168-174): `slots` = HashRate x the bulk's
// capacity entries set to -1, then rows are inserted in row order at
// Hash(key bytes) % slots, stepping one slot on (modulo slots) while occupied, at most
// 10 slots; an insert that finds no empty slot is silently dropped
// (com/cpp/cpu-runtime.hpp:80-91).  Probe walks the same slots from the key's home
// and stops at the first empty one (com/cpp/cpu-runtime.hpp:125-138): every slot
// between a kept row's home and its slot was occupied when it was inserted and stays
// so, hence every kept row is found and a dropped row never is, and of several kept
// rows with one key the earliest lies first on the walk.  The join therefore returns
// "the first kept row with this key", and only the set of kept rows is needed, not
// the table.
//
// In-order insertion is sequential, but rows whose 10-slot windows [home, home+9]
// cannot meet are independent.  Sorted by home slot, a gap of more than 9 between
// consecutive homes separates clusters that share no slot (the last cluster joins the
// first when its window wraps round to the first home).  So: sort (home, row) by
// home, number the clusters, sort rows by cluster (stable, so row order inside a
// cluster), and let one thread replay each cluster's inserts in row order against a
// slot-occupancy bitmap.  Clusters own disjoint slot ranges; a bit is claimed with an
// atomic OR whose old value says whether the slot was free.
#include <hip/hip_runtime.h>

#include <climits>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "kernels.h"
#include "refhash.h"

namespace svm {
namespace k {

namespace {

constexpr int kBlock = 256;
constexpr int kProbes = 10;  // com/cpp/cpu-runtime.hpp:82

inline unsigned grid_for(uint64_t n, unsigned cap = 65536) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g == 0) g = 1;
  return (unsigned)(g > cap ? cap : g);
}

// Hash of an int32 key's four bytes in memory order, each as a signed char
// (com/cpp/cpu-runtime.hpp:47-58).
__device__ __forceinline__ uint32_t hash_i32(int32_t v) {
  uint32_t h = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) h = h * 31u + (uint32_t)(int32_t)(signed char)(uint32_t)(v >> (8 * b));
  return h & 0x7FFFFFFFu;
}

__global__ void k_home(const int32_t* __restrict__ keys, uint32_t n, uint32_t slots, uint32_t* __restrict__ home,
                       uint32_t* __restrict__ row) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    home[r] = hash_i32(keys[r]) % slots;
    row[r] = r;
  }
}

// Home slots of keys of `size` bytes each (row r at col + r*size): Hash over all
// bytes, or for strings StrHash up to the first NUL (com/cpp/cpu-runtime.hpp:47-66).
__global__ void k_home_bytes(const char* __restrict__ col, int64_t size, int str, uint32_t n, uint32_t slots,
                             uint32_t* __restrict__ home, uint32_t* __restrict__ row) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const char* p = col + (uint64_t)r * (uint64_t)size;
    uint32_t h = 0;
    for (int64_t k = 0; k < size; ++k) {
      if (str && p[k] == '\0') break;
      h = h * 31u + (uint32_t)(int32_t)(signed char)p[k];
    }
    home[r] = (h & 0x7FFFFFFFu) % slots;
    row[r] = r;
  }
}

// head[i] = 1 where a new cluster starts in home order
__global__ void k_heads(const uint32_t* __restrict__ hs, uint32_t n, uint32_t* __restrict__ head) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    head[i] = (i == 0 || hs[i] - hs[i - 1] > (uint32_t)(kProbes - 1)) ? 1u : 0u;
}

// cluster of each row (numbered from 0 in home order); the first cluster joins the
// last when the last window wraps round to the first home
__global__ void k_cluster_of(const uint32_t* __restrict__ hs, const uint32_t* __restrict__ rs,
                             const uint32_t* __restrict__ incl, uint32_t n, uint32_t slots,
                             uint32_t* __restrict__ cluster) {
  const uint32_t ncl = incl[n - 1];
  const bool wrap = (uint64_t)hs[0] + slots - hs[n - 1] <= (uint64_t)(kProbes - 1);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t c = incl[i] - 1u;
    if (wrap && c == 0 && ncl > 1) c = ncl - 1u;
    cluster[rs[i]] = c;
  }
}

// One thread per cluster (the entry that starts it) replays its inserts in row order;
// with `risky`, it also marks every entry of a cluster that drops.
__global__ void k_replay(const uint32_t* __restrict__ cl, const uint32_t* __restrict__ rows,
                         const uint32_t* __restrict__ home, uint32_t n, uint32_t slots, uint32_t* occ,
                         uint8_t* __restrict__ kept, uint8_t* __restrict__ risky, int32_t* __restrict__ table,
                         unsigned long long* dropped) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t c = cl[i];
    if (i > 0 && cl[i - 1] == c) continue;
    unsigned long long lost = 0;
    uint32_t j = i;
    for (; j < n && cl[j] == c; ++j) {
      const uint32_t r = rows[j];
      uint32_t s = home[r];
      bool placed = false;
      for (int t = 0; t < kProbes; ++t) {
        const uint32_t bit = 1u << (s & 31u);
        if (!(atomicOr(&occ[s >> 5], bit) & bit)) {
          placed = true;
          if (table) table[s] = (int32_t)r;
          break;
        }
        s = s + 1u == slots ? 0u : s + 1u;
      }
      if (kept) kept[r] = placed ? 1 : 0;
      lost += placed ? 0 : 1;
    }
    if (lost) {
      atomicAdd(dropped, lost);
      if (risky)
        for (uint32_t e = i; e < j; ++e) risky[e] = 1;
    }
  }
}

// key and home of each selected entry, and where clusters start among them
__global__ void k_gather(const uint32_t* __restrict__ srow, const uint32_t* __restrict__ scl, uint32_t m,
                         const int32_t* __restrict__ keys, const uint32_t* __restrict__ home, int32_t* __restrict__ ekey,
                         uint32_t* __restrict__ ehome, uint8_t* __restrict__ head) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const uint32_t r = srow[e];
    ekey[e] = keys[r];
    ehome[e] = home[r];
    head[e] = (e == 0 || scl[e] != scl[e - 1]) ? 1 : 0;
  }
}

// start[nc] = m; stats[0] = most entries in a cluster, stats[1] = widest slot range
// (all ones for a cluster whose probes wrap past the last slot)
__global__ void k_cluster_stats(uint32_t* start, uint32_t nc, uint32_t m, const uint32_t* __restrict__ ehome,
                                uint32_t slots, uint32_t* stats) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const uint32_t e0 = start[c], e1 = c + 1 < nc ? start[c + 1] : m;
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    for (uint32_t e = e0; e < e1; ++e) {
      lo = min(lo, ehome[e]);
      hi = max(hi, ehome[e]);
    }
    atomicMax(&stats[0], e1 - e0);
    const bool wraps = (uint64_t)hi + (kProbes - 1) >= slots;
    atomicMax(&stats[1], wraps ? 0xFFFFFFFFu : hi - lo + (uint32_t)kProbes);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) start[nc] = m;
}

// bit (key - kmin) of every key in keys[0..n)
__global__ void k_key_bits(const int32_t* __restrict__ keys, uint32_t n, int32_t kmin, uint32_t range,
                           uint32_t* bits) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const uint32_t off = (uint32_t)(keys[r] - kmin);
    if (off < range) atomicOr(&bits[off >> 5], 1u << (off & 31u));
  }
}

// *flag = 1 when a cluster holds kProbes + 1 entries whose key bit is set
__global__ void k_cluster_hits(const uint32_t* __restrict__ start, uint32_t nc, const int32_t* __restrict__ key,
                               int32_t kmin, uint32_t range, const uint32_t* __restrict__ bits, uint32_t* flag) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    uint32_t hits = 0;
    for (uint32_t e = start[c]; e < start[c + 1]; ++e) {
      const uint32_t off = (uint32_t)(key[e] - kmin);
      hits += (off < range && ((bits[off >> 5] >> (off & 31u)) & 1u)) ? 1u : 0u;
      if (hits > (uint32_t)kProbes) {
        *flag = 1u;
        break;
      }
    }
  }
}

// Membership and clearing of a dimension object's entries: the first-row array
// (INT_MAX = no row) or the packed map (0 = no row).
struct FirstRows {
  int32_t* first;
  __device__ bool has(uint32_t off) const { return first[off] != INT_MAX; }
  __device__ void clear(uint32_t off) const { first[off] = INT_MAX; }
};
struct PackedSlots {
  uint32_t* words;
  unsigned wbits;
  uint32_t* bits;  // existence bits kept beside the map, or null
  __device__ bool has(uint32_t off) const {
    if (wbits == 32u) return words[off] != 0u;
    const unsigned per = 32u / wbits;
    return ((words[off / per] >> ((off % per) * wbits)) & ((1u << wbits) - 1u)) != 0u;
  }
  __device__ void clear(uint32_t off) const {
    if (bits) atomicAnd(&bits[off >> 5], ~(1u << (off & 31u)));
    if (wbits == 32u) { words[off] = 0u; return; }
    const unsigned per = 32u / wbits;
    atomicAnd(&words[off / per], ~(((1u << wbits) - 1u) << ((off % per) * wbits)));
  }
};

// Replays, in each dropping cluster where the object holds kProbes + 1 or more rows,
// the object's inserts (entries whose key has an entry in `map`) and clears the entry
// of every key whose insert drops.
template <typename Slots>
__global__ void k_clear_dropped(const uint32_t* __restrict__ start, uint32_t nc, const int32_t* __restrict__ key,
                                const uint32_t* __restrict__ home, int32_t kmin, uint32_t range, Slots map) {
  constexpr int kWords = kReplaySpan / 64;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const uint32_t e0 = start[c], e1 = start[c + 1];
    uint32_t members = 0, base = 0xFFFFFFFFu;
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t off = (uint32_t)(key[e] - kmin);
      if (off < range && map.has(off)) {
        ++members;
        base = min(base, home[e]);
      }
    }
    if (members <= (uint32_t)kProbes) continue;
    // each key occurs once in the root (unique keys), so clearing one member's entry
    // leaves the membership of the others unchanged
    unsigned long long occ[kWords] = {};
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t off = (uint32_t)(key[e] - kmin);
      if (off >= range || !map.has(off)) continue;
      uint32_t pos = home[e] - base;  // < kReplaySpan - kProbes + 1: the cluster spans <= kReplaySpan slots
      bool placed = false;
      for (int t = 0; t < kProbes; ++t, ++pos) {
        const unsigned long long bit = 1ull << (pos & 63u);
        if (!(occ[pos >> 6] & bit)) {
          occ[pos >> 6] |= bit;
          placed = true;
          break;
        }
      }
      if (!placed) map.clear(off);
    }
  }
}

int bits_of(uint64_t v) {
  int b = 1;
  while (b < 32 && (1ull << b) <= v) ++b;
  return b;
}

}  // namespace

namespace {

// The replay of inserts whose home slots are home[0..n) (row ids 0..n-1 in `row`).
uint64_t replay_homes(const int32_t* keys, DevBuf& home, DevBuf& row, uint64_t n, uint64_t slots, uint8_t* kept,
                      HtClusters* out, int32_t* table, hipStream_t s) {
  const uint32_t n32 = (uint32_t)n, sl = (uint32_t)slots;
  DevBuf hs(n * 4 + 4, s), rs(n * 4 + 4, s), incl(n * 4 + 4, s), cl(n * 4 + 4, s), cl2(n * 4 + 4, s),
      rows2(n * 4 + 4, s), occ(((uint64_t)sl + 31) / 32 * 4 + 4, s), drop(8, s), risky(out ? n + 16 : 16, s);
  // sort by home slot; rocPRIM's radix sort is stable, so equal homes stay in row order
  const int hbits = bits_of(slots - 1);
  const int cbits = bits_of(n);
  size_t t1 = 0, t2 = 0, t3 = 0;
  SVM_HIP(rocprim::radix_sort_pairs(nullptr, t1, home.as<uint32_t>(), hs.as<uint32_t>(), row.as<uint32_t>(),
                                    rs.as<uint32_t>(), n, 0, hbits, s));
  SVM_HIP(rocprim::inclusive_scan(nullptr, t2, hs.as<uint32_t>(), incl.as<uint32_t>(), n,
                                  rocprim::plus<uint32_t>(), s));
  SVM_HIP(rocprim::radix_sort_pairs(nullptr, t3, cl.as<uint32_t>(), cl2.as<uint32_t>(), row.as<uint32_t>(),
                                    rows2.as<uint32_t>(), n, 0, cbits, s));
  DevBuf tmp(std::max(t1, std::max(t2, t3)) + 16, s);
  size_t tb = tmp.bytes;
  SVM_HIP(rocprim::radix_sort_pairs(tmp.p, tb, home.as<uint32_t>(), hs.as<uint32_t>(), row.as<uint32_t>(),
                                    rs.as<uint32_t>(), n, 0, hbits, s));
  k_heads<<<grid_for(n), kBlock, 0, s>>>(hs.as<uint32_t>(), n32, cl.as<uint32_t>());
  hip_check(hipGetLastError(), "ht_analyse heads");
  tb = tmp.bytes;
  SVM_HIP(rocprim::inclusive_scan(tmp.p, tb, cl.as<uint32_t>(), incl.as<uint32_t>(), n, rocprim::plus<uint32_t>(),
                                  s));
  k_cluster_of<<<grid_for(n), kBlock, 0, s>>>(hs.as<uint32_t>(), rs.as<uint32_t>(), incl.as<uint32_t>(), n32, sl,
                                               cl.as<uint32_t>());
  hip_check(hipGetLastError(), "ht_analyse clusters");
  // rows by cluster (the sorts leave their inputs as they were, so `row` still holds
  // 0..n-1 in order and each cluster keeps row order)
  tb = tmp.bytes;
  SVM_HIP(rocprim::radix_sort_pairs(tmp.p, tb, cl.as<uint32_t>(), cl2.as<uint32_t>(), row.as<uint32_t>(),
                                    rows2.as<uint32_t>(), n, 0, cbits, s));
  fill_u32(occ.p, ((uint64_t)sl + 31) / 32, 0u, s);
  fill_u32(drop.p, 2, 0u, s);
  if (out) fill_u32(risky.p, (n + 3) / 4, 0u, s);
  if (table) fill_u32(table, sl, 0xFFFFFFFFu, s);
  k_replay<<<grid_for(n), kBlock, 0, s>>>(cl2.as<uint32_t>(), rows2.as<uint32_t>(), home.as<uint32_t>(), n32, sl,
                                          occ.as<uint32_t>(), kept, out ? risky.as<uint8_t>() : nullptr, table,
                                          drop.as<unsigned long long>());
  hip_check(hipGetLastError(), "ht_analyse replay");
  unsigned long long d = 0;
  SVM_HIP(hipMemcpyAsync(&d, drop.p, 8, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  if (!out) return d;
  out->dropped = d;
  if (!d) return d;
  // the entries of clusters that drop, still grouped by cluster in row order
  DevBuf srow(n * 4 + 4, s), scl(n * 4 + 4, s), cnt(16, s);
  size_t t4 = 0, t5 = 0;
  SVM_HIP(rocprim::select(nullptr, t4, rows2.as<uint32_t>(), risky.as<uint8_t>(), srow.as<uint32_t>(),
                          cnt.as<unsigned>(), n, s));
  SVM_HIP(rocprim::select(nullptr, t5, rocprim::counting_iterator<uint32_t>(0u), risky.as<uint8_t>(),
                          srow.as<uint32_t>(), cnt.as<unsigned>(), n, s));
  DevBuf tmp2(std::max(t4, t5) + 16, s);
  tb = tmp2.bytes;
  SVM_HIP(rocprim::select(tmp2.p, tb, rows2.as<uint32_t>(), risky.as<uint8_t>(), srow.as<uint32_t>(),
                          cnt.as<unsigned>(), n, s));
  tb = tmp2.bytes;
  SVM_HIP(rocprim::select(tmp2.p, tb, cl2.as<uint32_t>(), risky.as<uint8_t>(), scl.as<uint32_t>(),
                          cnt.as<unsigned>() + 1, n, s));
  unsigned m2[2] = {0, 0};
  SVM_HIP(hipMemcpyAsync(m2, cnt.p, 8, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  const uint32_t m = m2[0];
  out->entries = m;
  out->key = DevBuf((uint64_t)m * 4 + 4);
  out->home = DevBuf((uint64_t)m * 4 + 4);
  DevBuf head((uint64_t)m + 16, s);
  k_gather<<<grid_for(m), kBlock, 0, s>>>(srow.as<uint32_t>(), scl.as<uint32_t>(), m, keys, home.as<uint32_t>(),
                                          out->key.as<int32_t>(), out->home.as<uint32_t>(), head.as<uint8_t>());
  hip_check(hipGetLastError(), "ht_analyse gather");
  out->start = DevBuf((uint64_t)m * 4 + 8);
  tb = tmp2.bytes;
  SVM_HIP(rocprim::select(tmp2.p, tb, rocprim::counting_iterator<uint32_t>(0u), head.as<uint8_t>(),
                          out->start.as<uint32_t>(), cnt.as<unsigned>(), m, s));
  unsigned nc = 0;
  SVM_HIP(hipMemcpyAsync(&nc, cnt.p, 4, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  out->clusters = nc;
  DevBuf stats(16, s);
  fill_u32(stats.p, 4, 0u, s);
  k_cluster_stats<<<grid_for(nc), kBlock, 0, s>>>(out->start.as<uint32_t>(), nc, m, out->home.as<uint32_t>(), sl,
                                                  stats.as<uint32_t>());
  hip_check(hipGetLastError(), "ht_analyse stats");
  uint32_t st[2] = {0, 0};
  SVM_HIP(hipMemcpyAsync(st, stats.p, 8, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  out->max_rows = st[0];
  out->max_span = st[1];
  return d;
}

void check_sizes(uint64_t n, uint64_t slots) {
  if (slots == 0 || slots > 0xFFFFFFFFull) fail(SVM_ERR_INVALID, "hash table size out of range");
  if (n > 0x7FFFFFFFull) fail(SVM_ERR_UNSUPPORTED, "more than 2^31 rows in one bulk");
}

}  // namespace

uint64_t ht_analyse(const int32_t* keys, uint64_t n, uint64_t slots, uint8_t* kept, HtClusters* out, hipStream_t s) {
  if (out) *out = HtClusters();
  if (!n) return 0;
  check_sizes(n, slots);
  DevBuf home(n * 4 + 4, s), row(n * 4 + 4, s);
  k_home<<<grid_for(n), kBlock, 0, s>>>(keys, (uint32_t)n, (uint32_t)slots, home.as<uint32_t>(), row.as<uint32_t>());
  hip_check(hipGetLastError(), "ht_analyse home");
  return replay_homes(keys, home, row, n, slots, kept, out, nullptr, s);
}

uint64_t ht_build(const char* col, int64_t size, bool str, uint64_t n, uint64_t slots, int32_t* table,
                  hipStream_t s) {
  check_sizes(n, slots);
  if (!n) {
    fill_u32(table, slots, 0xFFFFFFFFu, s);
    return 0;
  }
  DevBuf home(n * 4 + 4, s), row(n * 4 + 4, s);
  k_home_bytes<<<grid_for(n), kBlock, 0, s>>>(col, size, str ? 1 : 0, (uint32_t)n, (uint32_t)slots,
                                              home.as<uint32_t>(), row.as<uint32_t>());
  hip_check(hipGetLastError(), "ht_build home");
  return replay_homes(nullptr, home, row, n, slots, nullptr, nullptr, table, s);
}

bool ht_subset_may_drop(const HtClusters& hc, const int32_t* keys, uint64_t n, int32_t kmin, uint64_t range,
                        hipStream_t s) {
  if (!hc.dropped || n <= (uint64_t)kProbes) return false;
  if (range == 0 || range > (1ull << 31) || n > 0x7FFFFFFFull) return true;
  const uint64_t words = (range + 31) / 32;
  DevBuf bits(words * 4 + 4, s), flag(8, s);
  fill_u32(bits.p, words, 0u, s);
  fill_u32(flag.p, 1, 0u, s);
  k_key_bits<<<grid_for(n), kBlock, 0, s>>>(keys, (uint32_t)n, kmin, (uint32_t)range, bits.as<uint32_t>());
  hip_check(hipGetLastError(), "ht_subset key bits");
  k_cluster_hits<<<grid_for(hc.clusters), kBlock, 0, s>>>(hc.start.as<uint32_t>(), hc.clusters, hc.key.as<int32_t>(),
                                                           kmin, (uint32_t)range, bits.as<uint32_t>(),
                                                           flag.as<uint32_t>());
  hip_check(hipGetLastError(), "ht_subset hits");
  uint32_t f = 0;
  SVM_HIP(hipMemcpyAsync(&f, flag.p, 4, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  return f != 0;
}

void ht_clear_dropped(const HtClusters& hc, int32_t kmin, uint32_t range, int32_t* first, hipStream_t s) {
  if (!hc.dropped || !hc.clusters) return;
  if (hc.max_span > (uint32_t)kReplaySpan) fail(SVM_ERR_INVALID, "ht_clear_dropped: cluster wider than the replay window");
  k_clear_dropped<FirstRows><<<grid_for(hc.clusters), kBlock, 0, s>>>(
      hc.start.as<uint32_t>(), hc.clusters, hc.key.as<int32_t>(), hc.home.as<uint32_t>(), kmin, range,
      FirstRows{first});
  hip_check(hipGetLastError(), "ht_clear_dropped");
}

void ht_clear_dropped_packed(const HtClusters& hc, int32_t kmin, uint32_t range, uint32_t* packed, int wbits,
                             uint32_t* bits, hipStream_t s) {
  if (!hc.dropped || !hc.clusters) return;
  if (hc.max_span > (uint32_t)kReplaySpan) fail(SVM_ERR_INVALID, "ht_clear_dropped: cluster wider than the replay window");
  k_clear_dropped<PackedSlots><<<grid_for(hc.clusters), kBlock, 0, s>>>(
      hc.start.as<uint32_t>(), hc.clusters, hc.key.as<int32_t>(), hc.home.as<uint32_t>(), kmin, range,
      PackedSlots{packed, (unsigned)wbits, bits});
  hip_check(hipGetLastError(), "ht_clear_dropped_packed");
}

uint64_t ht_kept(const int32_t* keys, uint64_t n, uint64_t slots, uint8_t* kept, hipStream_t s) {
  return ht_analyse(keys, n, slots, kept, nullptr, s);
}

}  // namespace k
}  // namespace svm
