// stats.cpp — cached per-column facts of resident tables (see stats.h).
//
// A column's key range, dictionary, kept rows and hash tables are computed on first
// use and reused by every later query.  Each entry holds a weak pointer to its table
// and the table's data_version() at the time: it is recomputed when the id is reused
// for a new table or when svm_table_put_column has rewritten the table's columns.
#include "stats.h"

#include <algorithm>
#include <cstdlib>

#include "kernels.h"

namespace svm {

namespace {
std::mutex g_cache_mu;
std::map<std::pair<const Table*, int>, std::shared_ptr<DictEntry>> g_dicts;
std::map<std::pair<const Table*, int>, KeyEntry> g_keys;

struct KeptEntry {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;
  uint64_t dropped = 0;
  DevBuf kept;  // one byte per row; empty when nothing is dropped
};
// (table, column, rows inserted, slots)
std::map<std::tuple<const Table*, int, uint64_t, uint64_t>, KeptEntry> g_kept;

struct TableEntry {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;
  uint64_t dropped = 0;
  DevBuf table;  // slots int32
};
std::map<std::tuple<const Table*, int, uint64_t, uint64_t>, TableEntry> g_tables;

// Entries of tables that no longer exist (a served query's dimension objects come
// and go) are removed once a cache grows past a few hundred entries.
template <typename M>
void prune_dead(M& m) {
  if (m.size() < 256) return;
  for (auto it = m.begin(); it != m.end();)
    it = it->second.owner.expired() ? m.erase(it) : std::next(it);
}

struct ClusterEntry {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;
  std::shared_ptr<const k::HtClusters> clusters;  // null: nothing dropped
};

// An entry computed for table t at its current contents.
template <typename E>
bool current(const E& e, const TablePtr& t) {
  return e.owner.lock() == t && e.version == t->data_version();
}
std::map<std::tuple<const Table*, int, uint64_t>, ClusterEntry> g_clusters;
std::map<std::pair<const Table*, int>, std::shared_ptr<const PackedKeys>> g_packed;

}  // namespace

namespace {
template <typename E>
bool dead(const E& e) { return e.owner.expired(); }
template <typename E>
bool dead(const std::shared_ptr<E>& e) { return !e || e->owner.expired(); }
template <typename M>
void sweep(M& m) {
  for (auto it = m.begin(); it != m.end();) it = dead(it->second) ? m.erase(it) : std::next(it);
}
}  // namespace

void release_dead_derived() {
  std::lock_guard<std::mutex> g(g_cache_mu);
  sweep(g_dicts);
  sweep(g_keys);
  sweep(g_kept);
  sweep(g_tables);
  sweep(g_clusters);
  sweep(g_packed);
}

std::shared_ptr<const PackedKeys> column_packed_keys(const TablePtr& t, int col, hipStream_t s) {
  if (t->base || t->row0 != 0 || col < 0 || col >= (int)t->cols.size() || t->cols[col].type != Int32 ||
      !t->cols[col].dev || !t->rows)
    return nullptr;
  const auto key = std::make_pair(t.get(), col);
  // a high plane of at least `forced` bits (a cached copy narrower than that is rebuilt)
  const char* fe = getenv("SVM_PACK_HIGH_BITS");
  const int forced = fe ? (atoi(fe) >= 8 ? 8 : atoi(fe) >= 2 ? 2 : 0) : 0;
  {
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto it = g_packed.find(key);
    if (it != g_packed.end() && current(*it->second, t) && it->second->high_bits >= forced) return it->second;
  }
  // the key range is cached: a column too wide is told apart without a kernel
  const KeyEntry ke = table_key(t, col, s);
  const int64_t span = (int64_t)ke.kmax - (int64_t)ke.kmin;
  if (span >= (int64_t)(1 << 24)) return nullptr;
  const int hb = std::max(span < (1 << 16) ? 0 : span < (1 << 18) ? 2 : 8, forced);
  noalloc_refuse("a narrow copy of a column");
  std::shared_ptr<PackedKeys> p;
  {
    p = std::make_shared<PackedKeys>();
    p->owner = t;
    p->version = t->data_version();
    p->kmin = ke.kmin;
    p->high_bits = hb;
    const uint64_t blocks = (t->rows + 511) / 512, bb = 1024 + 64 * (uint64_t)hb;
    p->bytes = DevBuf(blocks * bb + 256);
    // the last block and the padding are zeroed first (a partial block's unused rows)
    k::fill_u32(p->bytes.as<char>() + (blocks - 1) * bb, (bb + 256) / 4, 0u, s);
    k::pack_keys((const int32_t*)t->col_ptr(col), t->rows, ke.kmin, hb, p->bytes.as<uint8_t>(), s);
    SVM_HIP(hipStreamSynchronize(s));
  }
  std::lock_guard<std::mutex> g(g_cache_mu);
  if (g_packed.size() >= 64)
    for (auto it = g_packed.begin(); it != g_packed.end();)
      it = it->second->owner.expired() ? g_packed.erase(it) : std::next(it);
  g_packed[key] = p;
  return p;
}

std::pair<TablePtr, int> column_root(const TablePtr& t, int col) {
  TablePtr cur = t;
  int c = col;
  for (int hop = 0; hop < 16; ++hop) {
    if (c < 0 || c >= (int)cur->cols.size()) break;
    const Column& k = cur->cols[c];
    TablePtr up = k.origin.lock();
    // a projection holds a subset of its origin's values only while the origin still
    // holds the values it was projected from
    if (!up || k.origin_col < 0 || up->data_version() != k.origin_version) break;
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
  if (it != g_dicts.end() && current(*it->second, t)) return it->second;
  auto d = std::make_shared<DictEntry>();
  d->owner = t;
  d->version = t->data_version();
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
  if (it != g_keys.end() && current(it->second, t)) return it->second;
  KeyEntry e;
  e.owner = t;
  e.version = t->data_version();
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

const uint8_t* join_kept(const TablePtr& t, int col, uint64_t* dropped, hipStream_t s) {
  *dropped = 0;
  const uint64_t n = t->bulk0_rows();
  // an insert can only fail with 10 slots already taken: the first 10 always succeed
  if (n <= 10) return nullptr;
  if (t->cols[col].type != Int32 || !t->col_ptr(col)) fail(SVM_ERR_INVALID, "join key column is not a materialised Int32 column");
  // the reference passes HashRate * max_tuple_count_ as an unsigned int
  // (com/cpp/com.hpp:169, core/pipeline.cu:105-106)
  const uint64_t slots = (uint64_t)(uint32_t)(kHashRate * (t->max_tuple_count ? t->max_tuple_count : n));
  if (!slots) fail(SVM_ERR_UNSUPPORTED, "join hash table of '" + t->id + "' has 0 slots (HashRate x bulk capacity wraps)");
  const auto key = std::make_tuple(t.get(), col, n, slots);
  {
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto it = g_kept.find(key);
    if (it != g_kept.end() && current(it->second, t)) {
      *dropped = it->second.dropped;
      return it->second.dropped ? it->second.kept.as<uint8_t>() : nullptr;
    }
  }
  const int32_t* keys = (const int32_t*)t->col_ptr(col);
  // Inserting a subsequence of a row sequence keeps every slot it occupies occupied in
  // the full sequence at the same point, so it drops only inside the full sequence's
  // dropping clusters: a root column that drops nothing, or a subset that holds at
  // most 10 rows of each such cluster, drops nothing.
  bool replay = true;
  auto root = column_root(t, col);
  if ((root.first != t || root.second != col) && root.first->cols[root.second].type == Int32 &&
      root.first->col_ptr(root.second)) {
    auto hc = table_drop_clusters(root.first, root.second, slots, s);
    if (!hc) {
      replay = false;
    } else {
      KeyEntry ke = table_key(root.first, root.second, s);
      const uint64_t range = (uint64_t)((int64_t)ke.kmax - (int64_t)ke.kmin + 1);
      replay = k::ht_subset_may_drop(*hc, keys, n, ke.kmin, range, s);
    }
  }
  KeptEntry e;
  e.owner = t;
  e.version = t->data_version();
  if (replay) {
    DevBuf kept(n + 16);  // cached past this run: not from the stream's per-run arena
    e.dropped = k::ht_kept(keys, n, slots, kept.as<uint8_t>(), s);
    if (e.dropped) e.kept = std::move(kept);
  }
  std::lock_guard<std::mutex> g(g_cache_mu);
  auto it = g_kept.find(key);
  if (it == g_kept.end() || !current(it->second, t)) {
    prune_dead(g_kept);
    g_kept[key] = std::move(e);
    it = g_kept.find(key);
  }
  *dropped = it->second.dropped;
  return it->second.dropped ? it->second.kept.as<uint8_t>() : nullptr;
}

std::shared_ptr<const k::HtClusters> table_drop_clusters(const TablePtr& t, int col, uint64_t slots,
                                                         hipStream_t s) {
  if (t->rows <= 10) return nullptr;
  if (t->cols[col].type != Int32 || !t->col_ptr(col)) fail(SVM_ERR_INVALID, "join key column is not a materialised Int32 column");
  std::lock_guard<std::mutex> g(g_cache_mu);
  auto key = std::make_tuple(t.get(), col, slots);
  auto it = g_clusters.find(key);
  if (it != g_clusters.end() && current(it->second, t)) return it->second.clusters;
  ClusterEntry e;
  e.owner = t;
  e.version = t->data_version();
  auto hc = std::make_shared<k::HtClusters>();
  k::ht_analyse((const int32_t*)t->col_ptr(col), t->rows, slots, nullptr, hc.get(), s);
  if (hc->dropped) e.clusters = hc;
  prune_dead(g_clusters);
  g_clusters[key] = e;
  return e.clusters;
}

const int32_t* join_table(const TablePtr& t, int col, uint64_t* slots, uint64_t* dropped, hipStream_t s) {
  const Column& c = t->cols[col];
  if (!t->col_ptr(col)) fail(SVM_ERR_INVALID, "join key column of '" + t->id + "' is not materialised");
  if (c.type != Int16 && c.type != Int32 && c.type != Int64 && c.type != String)
    fail(SVM_ERR_UNSUPPORTED, "join keys must be Int16, Int32, Int64 or String columns");
  const uint64_t n = t->bulk0_rows();
  // the reference passes HashRate * max_tuple_count_ as an unsigned int
  // (com/cpp/com.hpp:169, core/pipeline.cu:105-124)
  *slots = (uint64_t)(uint32_t)(kHashRate * (t->max_tuple_count ? t->max_tuple_count : n));
  if (!*slots) fail(SVM_ERR_UNSUPPORTED, "join hash table of '" + t->id + "' has 0 slots (HashRate x bulk capacity wraps)");
  const auto key = std::make_tuple(t.get(), col, n, *slots);
  {
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto it = g_tables.find(key);
    if (it != g_tables.end() && current(it->second, t)) {
      *dropped = it->second.dropped;
      return it->second.table.as<int32_t>();
    }
  }
  TableEntry e;
  e.owner = t;
  e.version = t->data_version();
  e.table = DevBuf(*slots * 4 + 16);  // cached past this run: not from the stream's arena
  e.dropped = k::ht_build(t->col_ptr(col), c.size, c.type == String, n, *slots, e.table.as<int32_t>(), s);
  std::lock_guard<std::mutex> g(g_cache_mu);
  auto it = g_tables.find(key);
  if (it == g_tables.end() || !current(it->second, t)) {
    prune_dead(g_tables);
    g_tables[key] = std::move(e);
    it = g_tables.find(key);
  }
  *dropped = it->second.dropped;
  return it->second.table.as<int32_t>();
}

uint64_t pipeline_ht_slots(const Engine& E) { return (uint64_t)(uint32_t)(kHashRate * E.cfg.tuple_per_bulk); }

}  // namespace svm
