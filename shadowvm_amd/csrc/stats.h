// stats.h — per-column facts of resident tables, computed once and cached:
// key range and uniqueness (for dense join maps) and the sorted dictionary of
// group-key values with every row's code.
#pragma once

#include "refhash.h"
#include "svm_internal.h"

namespace svm {

struct DictEntry {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;  // owner's data_version() when computed
  DevBuf dict, codes;  // codes[r] for the rows of the owner table
  uint32_t n = 0;
};
struct KeyEntry {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;
  int32_t kmin = 0, kmax = -1;
  bool unique = true;
};

// The table and column a column's values were projected from (a select output's
// bare-column projection keeps the passing rows of its source column, in source
// order; a view is a row range of its base); the column itself when it has no such
// origin.  The column's rows are an in-order subsequence of the root's rows.
std::pair<TablePtr, int> column_root(const TablePtr& t, int col);

std::shared_ptr<DictEntry> table_dict(const TablePtr& t, int col, hipStream_t s);
KeyEntry table_key(const TablePtr& t, int col, hipStream_t s);

// A narrow copy of Int32 column `col` of `t` (a table that owns its columns, row 0
// at the start of the allocation), in blocks of 512 rows: row i's value - kmin has its
// low 16 bits at uint16 i % 512 of block i / 512 and its `high_bits` high bits (0, 2
// or 8: the fewest that hold the column's range) in a plane after them, row i % 512 at
// bit (i % 512) * high_bits of byte offset 1024 (1,024 + 64 x high_bits bytes per
// block: 16, 18 or 24 bits a row).  The compacting scans stream their tier-0 key
// column from it.  Null when the column's range needs more than 24 bits.  Built on
// first use, cached with the table.  SVM_PACK_HIGH_BITS=2 or 8 widens the high plane
// to at least that many bits.
struct PackedKeys {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;
  int32_t kmin = 0;
  int high_bits = 8;
  DevBuf bytes;
};
std::shared_ptr<const PackedKeys> column_packed_keys(const TablePtr& t, int col, hipStream_t s);

// Frees the cached facts (and their device memory: narrow key copies, kept-row bytes, hash
// tables) of tables that no longer exist.  Called when a client drops a table.
void release_dead_derived();

constexpr uint64_t kHashRate = 5;  // com/cpp/define.hpp:29

// Rows of bulk 0 of `t` that the reference's join hash table on Int32 column `col`
// keeps (HashRate x max_tuple_count slots, rows inserted in row order, at most 10
// probes, core/pipeline.cu:88-133): nullptr when it keeps every row, else one byte per
// row (1 kept, 0 dropped) on the device, valid while `t` lives.  *dropped receives the
// number of dropped inserts.  Cached per table.
const uint8_t* join_kept(const TablePtr& t, int col, uint64_t* dropped, hipStream_t s);

// The clusters of the reference's hash table over every row of Int32 column `col` of
// `t` (row order, `slots` slots) that drop an insert, or nullptr when it drops
// nothing (then no in-order subset of those rows drops anything either).  Cached.
std::shared_ptr<const k::HtClusters> table_drop_clusters(const TablePtr& t, int col, uint64_t slots,
                                                         hipStream_t s);

// The reference's hash table over bulk 0 of column `col` of `t` (Int16/Int32/Int64
// keys hashed over their bytes, String keys with StrHash), HashRate x max_tuple_count
// slots: table[slot] = row or -1, valid while `t` lives.  *slots and *dropped receive
// its size and the inserts it dropped.  Cached per table.
const int32_t* join_table(const TablePtr& t, int col, uint64_t* slots, uint64_t* dropped, hipStream_t s);

// Slots of the hash table on a dimension object made by a pipeline (a bulk of
// tuple_per_bulk rows): HashRate x tuple_per_bulk, as the reference's unsigned int.
uint64_t pipeline_ht_slots(const Engine& E);

}  // namespace svm
