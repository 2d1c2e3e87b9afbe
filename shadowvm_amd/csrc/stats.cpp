// stats.cpp — cached per-column facts of resident tables (see stats.h).
//
// Tables in the store are immutable once loaded, so a column's key range and its
// dictionary are computed on first use and reused by every later query; the cache
// holds weak pointers and recomputes when a table id is reused for a new table.
#include "stats.h"

#include "kernels.h"

namespace svm {

namespace {
std::mutex g_cache_mu;
std::map<std::pair<const Table*, int>, std::shared_ptr<DictEntry>> g_dicts;
std::map<std::pair<const Table*, int>, KeyEntry> g_keys;
}  // namespace

std::pair<TablePtr, int> column_root(const TablePtr& t, int col) {
  TablePtr cur = t;
  int c = col;
  for (int hop = 0; hop < 16; ++hop) {
    if (c < 0 || c >= (int)cur->cols.size()) break;
    const Column& k = cur->cols[c];
    TablePtr up = k.origin.lock();
    if (!up || k.origin_col < 0) break;
    cur = up;
    c = k.origin_col;
  }
  if (cur->base) cur = cur->base;  // a view's values are a subset of its base's
  return {cur, c};
}

std::shared_ptr<DictEntry> table_dict(const TablePtr& t, int col, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  auto key = std::make_pair(t.get(), col);
  auto it = g_dicts.find(key);
  if (it != g_dicts.end() && it->second->owner.lock() == t) return it->second;
  auto d = std::make_shared<DictEntry>();
  d->owner = t;
  uint64_t n = t->rows;
  DevBuf vals(n * 4 + 4);
  k::values_of(t->col_ptr(col), t->cols[col].type, t->cols[col].size, n, vals.as<int32_t>(), s);
  d->dict = DevBuf(n * 4 + 4);
  d->n = (uint32_t)k::dict_build(vals.as<int32_t>(), n, d->dict.as<int32_t>(), s);
  d->codes = DevBuf(n * 4 + 4);
  k::codes_of(vals.as<int32_t>(), n, d->dict.as<int32_t>(), d->n, d->codes.as<int32_t>(), s);
  SVM_HIP(hipStreamSynchronize(s));
  g_dicts[key] = d;
  return d;
}

KeyEntry table_key(const TablePtr& t, int col, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  auto key = std::make_pair(t.get(), col);
  auto it = g_keys.find(key);
  if (it != g_keys.end() && it->second.owner.lock() == t) return it->second;
  KeyEntry e;
  e.owner = t;
  if (t->rows) {
    DevBuf mm(8);
    k::minmax_i32((const int32_t*)t->col_ptr(col), t->rows, mm.as<int32_t>(), s);
    int32_t h[2];
    SVM_HIP(hipMemcpyAsync(h, mm.p, 8, hipMemcpyDeviceToHost, s));
    SVM_HIP(hipStreamSynchronize(s));
    e.kmin = h[0];
    e.kmax = h[1];
    uint64_t range = (uint64_t)((int64_t)h[1] - (int64_t)h[0] + 1);
    e.unique = range <= (1ull << 28) &&
               !k::has_duplicates((const int32_t*)t->col_ptr(col), t->rows, h[0], (uint32_t)range, s);
  }
  g_keys[key] = e;
  return e;
}

}  // namespace svm
