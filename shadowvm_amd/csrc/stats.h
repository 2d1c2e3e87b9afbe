// stats.h — per-column facts of resident tables, computed once and cached:
// key range and uniqueness (for dense join maps) and the sorted dictionary of
// group-key values with every row's code.
#pragma once

#include "svm_internal.h"

namespace svm {

struct DictEntry {
  std::weak_ptr<Table> owner;
  DevBuf dict, codes;  // codes[r] for the rows of the owner table
  uint32_t n = 0;
};
struct KeyEntry {
  std::weak_ptr<Table> owner;
  int32_t kmin = 0, kmax = -1;
  bool unique = true;
};

// The table and column a column's values were projected from (a select output's
// bare-column projection keeps a subset of its source's values); the column itself
// when it has no such origin.
std::pair<TablePtr, int> column_root(const TablePtr& t, int col);

std::shared_ptr<DictEntry> table_dict(const TablePtr& t, int col, hipStream_t s);
KeyEntry table_key(const TablePtr& t, int col, hipStream_t s);

}  // namespace svm
