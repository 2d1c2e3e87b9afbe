// refhash.h — the reference's join hash table, reduced to which rows it keeps.
#pragma once

#include "svm_internal.h"

namespace svm {
namespace k {

// The clusters of a table's inserts that drop something (only there can a subset of
// its rows, inserted in the same order, drop anything).  Entries are grouped by
// cluster, in row order inside a cluster: entry e has the key key[e] and the home slot
// home[e]; cluster c holds entries [start[c], start[c+1]).
struct HtClusters {
  uint64_t dropped = 0;     // inserts the full table drops
  uint32_t clusters = 0;    // clusters with a drop
  uint32_t entries = 0;     // their inserts
  uint32_t max_rows = 0;    // most inserts in one of them
  uint32_t max_span = 0;    // widest slot range of one of them (last home - first home + 10;
                            // all ones when a cluster's probes wrap past the last slot)
  DevBuf start, key, home;  // uint32 [clusters + 1], int32 [entries], uint32 [entries]
};

// The reference's hash table over keys[0..n), inserted in row order into `slots` slots
// (Hash of the key's bytes % slots, at most 10 linear probes, silent drop;
// com/cpp/cpu-runtime.hpp:80-91).  Writes kept[r] = 1 if row r's insert found a slot
// (kept may be null) and, when `clusters` is given, the clusters that drop.  Returns
// the number of dropped inserts (synchronises).
uint64_t ht_analyse(const int32_t* keys, uint64_t n, uint64_t slots, uint8_t* kept, HtClusters* clusters,
                    hipStream_t s);

// False when an in-order subset of the rows behind `hc` (the subset's keys are
// keys[0..n), all within [kmin, kmin + range)) certainly drops nothing: a subset drops
// only inside a cluster that drops, and only where it holds at least 11 of the
// cluster's inserts (a drop needs 10 slots taken by earlier inserts of the cluster).
// Inserts are matched by key, so duplicate keys only make the answer more cautious.
// Synchronises.
bool ht_subset_may_drop(const HtClusters& hc, const int32_t* keys, uint64_t n, int32_t kmin, uint64_t range,
                        hipStream_t s);

// The reference's hash table itself (BuildHashTable, com/cpp/com.hpp:168-174) over n
// keys of `size` bytes each (row r at col + r*size), hashed with Hash over all bytes or,
// for strings, StrHash up to the first NUL (InsertHashTable<char>,
// com/cpp/cpu-runtime.hpp:92-103): table[slot] = row, -1 for an empty slot.  Returns
// the number of dropped inserts (synchronises).
uint64_t ht_build(const char* col, int64_t size, bool str, uint64_t n, uint64_t slots, int32_t* table,
                  hipStream_t s);

// Widest cluster (in slots) replayed in registers by the subset kernels.
constexpr int kReplaySpan = 256;

// For a dimension object that is an in-order subset, with unique keys, of the rows
// behind `hc`, whose dense map first[key - kmin] (range entries) holds the object's row
// of each key or INT32_MAX: clears the entries of the keys whose inserts the
// reference's table over the object's rows drops.  Replays only the clusters where
// the object holds at least 11 rows; needs hc.max_span <= kReplaySpan.  No
// synchronisation.
void ht_clear_dropped(const HtClusters& hc, int32_t kmin, uint32_t range, int32_t* first, hipStream_t s);
// The same over a packed map (an entry is cleared to 0, and its existence bit when
// `bits` is given).
void ht_clear_dropped_packed(const HtClusters& hc, int32_t kmin, uint32_t range, uint32_t* packed, int wbits,
                             uint32_t* bits, hipStream_t s);

}  // namespace k
}  // namespace svm
