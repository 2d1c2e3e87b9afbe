// exec.cpp — one pipeline on the GPU: resolve inputs, build dimension maps,
// generate + launch the fused kernel, emit the output table.
//
// Reference flow being replaced (per pipeline):
//   Pipeline::Open   host BuildHashTable per join            core/pipeline.cu:82-137
//   Pipeline::Run    one thread + kernel per 2M-row bulk     core/pipeline.cu:139-254
//   TransformTask    Query::Next, ShrinkTask, ReduceTask     core/pipeline.cu:262-472
// Here: Open becomes a GPU dense-map build (plus per-key dictionaries), Run is a
// single persistent scan over the HBM-resident source, and the reduce happens in
// the scan (LDS tables + 64-bit atomics) followed by an ordered group emission.
#include "exec.h"

#include "loader.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <set>

#include "codegen.h"
#include "host_engine.h"
#include "jit.h"
#include "kargs.h"
#include "kernels.h"
#include "stats.h"

namespace svm {

namespace {

constexpr uint64_t kMaxDenseRange = 1ull << 28;  // keys per dense join map
// Beyond this many keys a dense map must hold at least one dimension row per
// kMaxDenseSpread keys of its range, else the join probes the hash table.
constexpr uint64_t kSparseDenseFloor = 1ull << 24;
constexpr uint64_t kMaxDenseSpread = 64;
constexpr uint64_t kMaxGroups = 1ull << 27;
constexpr uint64_t kMaxFactKeyRange = 1ull << 24;
constexpr uint64_t kDenseAggBytes = 16 * 1024;  // LDS for a dense per-workgroup table
constexpr int kHashBits = 10;                    // LDS hash table slots = 2^kHashBits
constexpr uint64_t kLdsBudget = 64 * 1024;       // dynamic LDS per workgroup
// Largest output (source rows x projected bytes) a select allocates for every source
// row so that it can compact in one launch.
constexpr uint64_t kSingleSelectBytes = 256ull << 20;
// ... and at most this many 1024-row tiles: beyond a few hundred, the tiles' polling
// of their predecessors' statuses costs more than the second pass over the source.
constexpr uint64_t kSingleSelectTiles = 256;
// Rows above which a dense map with sub-byte slots is built by key, not by row.
constexpr uint64_t kDirectNarrowRows = 1ull << 16;
// Code maps larger than this are probed through an existence bitmap first (the scan
// reads the code only for rows that pass every join), so the probe chain touches
// 1 bit per key instead of the code width.
constexpr uint64_t kDeferMapBytes = 128 * 1024;

const char* ctype(int t) {
  switch (t) {
    case Byte: return "char";
    case Int16: return "short";
    case Int32: return "int";
    case Int64: return "long long";
    default: return nullptr;
  }
}

// Timing events are reused across runs (a run synchronises its stream before it
// returns, so an event handed back is idle; re-recording one is allowed anyway).
std::mutex g_event_mu;
std::vector<hipEvent_t> g_event_pool;
// Inside a no-allocation run an event is taken only from the pool (none is created:
// the run then goes untimed).
struct Event {
  hipEvent_t e = nullptr;
  Event() {
    {
      std::lock_guard<std::mutex> g(g_event_mu);
      if (!g_event_pool.empty()) {
        e = g_event_pool.back();
        g_event_pool.pop_back();
        return;
      }
    }
    if (!noalloc_active()) SVM_HIP(hipEventCreate(&e));
  }
  void record(hipStream_t s) const {
    if (e) SVM_HIP(hipEventRecord(e, s));
  }
  ~Event() {
    if (!e) return;
    std::lock_guard<std::mutex> g(g_event_mu);
    g_event_pool.push_back(e);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
};

// The first `ncols` columns of an output table copied to host vectors once the work
// on stream `s` is done (this synchronises `s`): through the stream's pinned staging
// buffer with one synchronisation when they fit in kStagedCopyBytes, else one
// synchronous copy per column.
constexpr uint64_t kStagedCopyBytes = 64ull << 20;
// Group spaces whose every-group-id output fits this many bytes per column set are
// emitted with one synchronisation (the output and a host-mapped copy sized by G).
constexpr uint64_t kOneSyncEmitBytes = 16ull << 20;
uint64_t row_bytes_of(const std::vector<std::pair<int, int64_t>>& schema) {
  uint64_t b = 0;
  for (auto& c : schema) b += (uint64_t)c.second;
  return b;
}
void copy_out(const Table& out, const std::vector<std::pair<int, int64_t>>& schema, int ncols, uint64_t rows,
              hipStream_t s, std::vector<std::vector<uint8_t>>& cols) {
  uint64_t total = 0;
  for (int q = 0; q < ncols; ++q) total += rows * (uint64_t)schema[q].second;
  cols.clear();
  if (total && total <= kStagedCopyBytes) {
    char* st = (char*)host_staging(s, total);
    uint64_t off = 0;
    for (int q = 0; q < ncols; ++q) {
      const uint64_t b = rows * (uint64_t)schema[q].second;
      if (b) SVM_HIP(hipMemcpyAsync(st + off, out.col_ptr(q), b, hipMemcpyDeviceToHost, s));
      off += b;
    }
    SVM_HIP(hipStreamSynchronize(s));
    off = 0;
    for (int q = 0; q < ncols; ++q) {
      const uint64_t b = rows * (uint64_t)schema[q].second;
      cols.emplace_back(st + off, st + off + b);
      off += b;
    }
    return;
  }
  SVM_HIP(hipStreamSynchronize(s));  // the copies below are not ordered after the stream's work
  for (int q = 0; q < ncols; ++q) {
    std::vector<uint8_t> h(rows * (uint64_t)schema[q].second);
    if (!h.empty()) SVM_HIP(hipMemcpy(h.data(), out.col_ptr(q), h.size(), hipMemcpyDeviceToHost));
    cols.push_back(std::move(h));
  }
}

// The source column an expression names when it is exactly "$(0.c)" and the output
// column has that column's type and width (the projection then copies it unchanged);
// -1 otherwise.
int bare_source_col(const std::string& e, const Schemas& schemas, int type, int64_t size) {
  int obj = -1, col = -1;
  const size_t b = e.find_first_not_of(" \t"), t = e.find_last_not_of(" \t");
  if (b == std::string::npos) return -1;
  const std::string x = e.substr(b, t - b + 1);
  if (sscanf(x.c_str(), "$(%d.%d)", &obj, &col) != 2 || obj != 0) return -1;
  if (x != "$(0." + std::to_string(col) + ")") return -1;
  if (col < 0 || col >= (int)schemas[0].size()) return -1;
  if (schemas[0][col].first != type || schemas[0][col].second != size) return -1;
  return col;
}

double host_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

float elapsed(const Event& a, const Event& b) {
  float ms = 0;
  if (a.e && b.e) SVM_HIP(hipEventElapsedTime(&ms, a.e, b.e));
  return ms;
}

struct JoinBuild {
  JoinShape shape;
  int rcol = 0;
  int32_t kmin = 0;
  uint32_t range = 0;
  uint64_t words = 0;
  DevBuf packed;
  uint32_t lds_off = 0;
  // LDS placement for staged scans (assign_lds): whether the map, or a deferred map's
  // existence bits, is staged and at which word
  bool staged_lds = false, staged_bits = false;
  uint32_t staged_off = 0;
  uint64_t dropped = 0;  // inserts the reference's hash table drops
  const int32_t* table = nullptr;  // probe joins: the reference's hash table (slots entries)
  uint64_t slots = 0;
  // estimated fraction of source rows whose key the map holds: the dimension object's
  // rows over its key column's root table's (foreign keys assumed spread over the
  // root's keys); 1 when the object is not drawn from another table
  double sel = 1.0;
};

struct KeyBuild {
  KeyShape shape;
  uint32_t n = 0, stride = 1;
  uint64_t card = 0;  // codes 0..card-1 (a source-column key's range may reach 2^32)
  int bits = 0, shift = 0;  // its field in a packed 64-bit key (wide group spaces)
  int32_t fmin = 0;
  std::shared_ptr<DictEntry> dict;  // sorted values of the key column's origin
  DevBuf codes;                     // code of every dimension row (row-id payloads)
};

int width_for(uint64_t maxval) {
  const int ws[] = {1, 2, 4, 8, 16, 32};
  for (int w : ws)
    if (w == 32 || maxval <= ((1ull << w) - 1)) return w;
  return 32;
}

// Group-key component of a dimension: the dictionary of the column's origin (the
// base table it was projected from, cached across queries, so no host round trip).
void key_dict(const TablePtr& D, int col, KeyBuild& kb, hipStream_t s) {
  const Column& c = D->cols[col];
  if (!c.dev) fail(SVM_ERR_INVALID, "group key column is not materialised");
  if (c.type != Int32 && c.type != String && c.type != Int16 && c.type != Byte)
    fail(SVM_ERR_UNSUPPORTED, "group keys must be integer or string columns");
  auto root = column_root(D, col);
  kb.dict = table_dict(root.first, root.second, s);
  kb.n = kb.dict->n;
}

// ... and, by binary search, the code of every row of this dimension object.
void build_key_dict(const TablePtr& D, int col, KeyBuild& kb, DevBuf* codes_out, hipStream_t s) {
  key_dict(D, col, kb, s);
  const Column& c = D->cols[col];
  uint64_t n = D->rows;
  DevBuf vals(n * 4 + 4, s);
  k::values_of(D->col_ptr(col), c.type, c.size, n, vals.as<int32_t>(), s);
  *codes_out = DevBuf(n * 4 + 4, s);
  k::codes_of(vals.as<int32_t>(), n, kb.dict->dict.as<int32_t>(), kb.n, codes_out->as<int32_t>(), s);
}

struct Ctx {
  Engine& E;
  const Plan& plan;
  const PipelineInfo& info;
  hipStream_t s;
  TablePtr src;
  std::vector<TablePtr> refs;
  Schemas schemas;
};

void resolve(Ctx& c) {
  if (c.info.pre_pipeline_id.empty())
    fail(SVM_ERR_NOT_FOUND, "pipeline '" + c.info.id + "' names no source object (pre_pipeline_id[0])");
  c.src = c.E.store.get(c.info.pre_pipeline_id[0]);
  for (size_t k = 1; k < c.info.pre_pipeline_id.size(); ++k)
    c.refs.push_back(c.E.store.get(c.info.pre_pipeline_id[k]));
  if (c.refs.size() + 1 > SVM_MAX_JOIN)
    fail(SVM_ERR_UNSUPPORTED, "at most 7 ref objects (join_index[8], codegen/template.cpp:65)");
  std::vector<std::pair<int, int64_t>> s0;
  for (auto& col : c.src->cols) s0.push_back({col.type, col.size});
  const auto& decl = c.plan.source_info.column_list;
  if (!decl.empty()) {
    if (decl.size() != s0.size())
      fail(SVM_ERR_INVALID, "Source schema has " + std::to_string(decl.size()) + " columns, object '" +
                                c.src->id + "' has " + std::to_string(s0.size()));
    for (size_t q = 0; q < decl.size(); ++q)
      if (decl[q].type != s0[q].first || (decl[q].val_size && decl[q].val_size != s0[q].second))
        fail(SVM_ERR_INVALID, "Source column " + std::to_string(q) + " does not match object '" + c.src->id + "'");
  }
  c.schemas.push_back(s0);
  for (auto& r : c.refs) {
    std::vector<std::pair<int, int64_t>> sr;
    for (auto& col : r->cols) sr.push_back({col.type, col.size});
    c.schemas.push_back(sr);
  }
  if (!c.plan.has_sink) fail(SVM_ERR_INVALID, "pipeline '" + c.info.id + "' has no Sink");
}

// Joins: validation and dense-map builds.  `exprs_refs` are (obj,col) operands read
// by metric / projection expressions (those dimensions need row-id payloads).
std::vector<JoinBuild> build_joins(Ctx& c, const std::set<std::pair<int, int>>& expr_refs,
                                   std::vector<KeyBuild>& keys, bool allow_defer) {
  std::vector<JoinBuild> joins;
  std::set<int> seen;
  for (size_t j = 0; j < c.plan.join_rkey.size(); ++j) {
    KeyIndex L = c.plan.join_lkey[j], R = c.plan.join_rkey[j];
    if (L.obj != 0) fail(SVM_ERR_UNSUPPORTED, "JoinWith lkey must be a source column");
    if (R.obj < 1 || R.obj > (int)c.refs.size())
      fail(SVM_ERR_INVALID, "JoinWith rkey names ref object " + std::to_string(R.obj) + " which is not an input");
    if (!seen.insert(R.obj).second) fail(SVM_ERR_UNSUPPORTED, "two JoinWith on the same ref object");
    if (L.col >= (int)c.schemas[0].size() || R.col >= (int)c.schemas[R.obj].size())
      fail(SVM_ERR_INVALID, "JoinWith key column out of range");
    const auto lt = c.schemas[0][L.col], rt = c.schemas[R.obj][R.col];
    // BuildHashTable is instantiated for these key types (core/pipeline.cu:97-129;
    // Byte keys get no table there)
    if (rt.first != Int16 && rt.first != Int32 && rt.first != Int64 && rt.first != String)
      fail(SVM_ERR_UNSUPPORTED, "join keys must be Int16, Int32, Int64 or String columns");
    // Probe reads the lkey at the rkey's type and width (codegen/codegen.cpp:137-141)
    if (lt != rt) fail(SVM_ERR_UNSUPPORTED, "JoinWith lkey and rkey must have the same type and width");
    const Table& D = *c.refs[R.obj - 1];
    const TablePtr& Dp = c.refs[R.obj - 1];
    // only bulk 0 of the dimension is joined (core/pipeline.cu:91,294-295)
    const uint64_t n0 = D.bulk0_rows();
    JoinBuild jb;
    jb.rcol = R.col;
    jb.shape.robj = R.obj;
    jb.shape.lcol = L.col;
    // Int32 keys within the dense-map range join through a dense map; other keys probe
    // the reference's hash table itself
    auto root = column_root(Dp, R.col);
    KeyEntry ke;
    uint64_t range = 0;
    if (rt.first == Int32 && n0) {
      ke = table_key(root.first, root.second, c.s);
      range = (uint64_t)((int64_t)ke.kmax - (int64_t)ke.kmin + 1);
    }
    // (a sparse key range too: a dense map wider than 2^24 slots holding fewer than one
    // key in 64 would cost far more memory than the hash table's kHashRate slots a row)
    const bool probe = rt.first != Int32 || range > kMaxDenseRange ||
                       (range > kSparseDenseFloor && range / kMaxDenseSpread > (n0 ? n0 : 1));
    if ((root.first != Dp || root.second != R.col) && root.first->bulk0_rows() > 0)
      jb.sel = std::min(1.0, (double)n0 / (double)root.first->bulk0_rows());
    // payload selection
    std::vector<int> comps;
    for (size_t q = 0; q < keys.size(); ++q)
      if (keys[q].shape.src != KEY_FACT && keys[q].shape.obj == R.obj) comps.push_back((int)q);
    bool refs = false;
    for (auto& o : expr_refs) refs = refs || o.first == R.obj;
    int payload = refs || comps.size() > 1 || probe ? PAY_ROW : comps.size() == 1 ? PAY_CODE : PAY_EXIST;
    // A dimension object drawn (in order) from a root table with unique keys has
    // unique keys itself: its dense map is built in one pass over its rows (each row
    // writes its own slot), and the inserts the reference's table would drop are
    // replayed inside the root's dropping clusters afterwards.
    const bool derived = root.first != Dp || root.second != R.col;
    const uint64_t slots = (uint64_t)(uint32_t)(kHashRate * (D.max_tuple_count ? D.max_tuple_count : n0));
    std::shared_ptr<const k::HtClusters> hc;
    bool direct = false;
    if (!probe && range && derived && ke.unique && slots) {
      hc = table_drop_clusters(root.first, root.second, slots, c.s);
      direct = !hc || hc->max_span <= (uint32_t)k::kReplaySpan;
    }
    // key dictionaries of this dimension (a code payload of the direct build is
    // looked up inside it)
    DevBuf codes_for_pack;
    const KeyBuild* direct_code = nullptr;
    for (int q : comps) {
      KeyBuild& kb = keys[q];
      kb.shape.join = (int)j;
      kb.shape.src = payload == PAY_CODE ? KEY_DIMCODE : KEY_DIMROW;
      if (direct && payload == PAY_CODE) {
        key_dict(Dp, kb.shape.col, kb, c.s);
        direct_code = &kb;
      } else {
        build_key_dict(Dp, kb.shape.col, kb, payload == PAY_CODE ? &codes_for_pack : &kb.codes, c.s);
      }
    }
    jb.shape.payload = payload;
    uint64_t maxval = payload == PAY_EXIST ? 1 : payload == PAY_CODE ? keys[comps[0]].n : D.rows;
    jb.shape.wbits = payload == PAY_ROW ? 32 : width_for(maxval);
    // slots narrower than a byte share words: many rows with neighbouring keys would
    // serialise on their atomicOr, so large objects build those maps by key instead
    if (direct && n0 > kDirectNarrowRows && jb.shape.wbits < 8) {
      direct = false;
      if (direct_code) {
        KeyBuild& kb = keys[comps[0]];
        build_key_dict(Dp, kb.shape.col, kb, &codes_for_pack, c.s);
        direct_code = nullptr;
      }
    }
    const void* kcol = D.col_ptr(R.col);
    if (!kcol) fail(SVM_ERR_INVALID, "join key column of '" + D.id + "' is not materialised");
    if (probe) {
      jb.shape.kind = rt.first == String ? JOIN_PROBE_STR : JOIN_PROBE;
      jb.shape.rcol = R.col;
      jb.shape.ktype = rt.first;
      jb.shape.ksize = (int)rt.second;
      jb.table = join_table(Dp, R.col, &jb.slots, &jb.dropped, c.s);
      if (jb.dropped && c.E.cfg.verbose)
        fprintf(stderr, "[svm] %s: the hash table on '%s' drops %llu of %llu inserts\n", c.info.id.c_str(),
                D.id.c_str(), (unsigned long long)jb.dropped, (unsigned long long)n0);
      joins.push_back(std::move(jb));
      continue;
    }
    // dense map over the key range of the column's origin (a superset of this
    // dimension object's keys, cached per resident table: no host round trip)
    if (n0) {
      jb.kmin = ke.kmin;
      jb.range = (uint32_t)range;
    }
    uint64_t per = 32 / (uint64_t)jb.shape.wbits;
    jb.words = ((uint64_t)jb.range + per - 1) / per;
    // a large code map is probed through its existence bits (stored after it)
    jb.shape.defer = allow_defer && payload == PAY_CODE && jb.words * 4 > kDeferMapBytes;
    const uint64_t bitwords = jb.shape.defer ? k::exist_words(jb.range) : 0;
    jb.packed = DevBuf((jb.words + bitwords) * 4 + 16, c.s);
    auto finish_map = [&](JoinBuild& b) {
      if (b.shape.defer)
        k::exist_bits(b.packed.as<uint32_t>(), b.shape.wbits, b.range, b.packed.as<uint32_t>() + b.words, c.s);
    };
    if (jb.range) {
      // The reference's table drops inserts that find no free slot in 10 probes; a
      // dropped row is never found, so the map takes the first kept row of each key.
      // A dimension object drawn (in order) from a root table with unique keys drops
      // only inside the root's dropping clusters: its map is built, then those
      // clusters are replayed on the device.  Otherwise the kept rows are computed
      // (and cached) for the object itself.
      if (direct) {
        const Column* ac = direct_code ? &D.cols[direct_code->shape.col] : nullptr;
        uint32_t* bits = jb.shape.defer ? jb.packed.as<uint32_t>() + jb.words : nullptr;
        k::pack_direct((const int32_t*)kcol, n0, jb.kmin, jb.range, payload,
                       ac ? D.col_ptr(direct_code->shape.col) : nullptr, ac ? ac->type : Int32, ac ? ac->size : 4,
                       direct_code ? direct_code->dict->dict.as<int32_t>() : nullptr, direct_code ? direct_code->n : 0,
                       jb.shape.wbits, jb.packed.as<uint32_t>(), bits, c.s);
        if (hc && n0 > 10)
          k::ht_clear_dropped_packed(*hc, jb.kmin, jb.range, jb.packed.as<uint32_t>(), jb.shape.wbits, bits, c.s);
        joins.push_back(std::move(jb));
        continue;
      }
      DevBuf first((uint64_t)jb.range * 4, c.s);
      {
        const uint8_t* kept = join_kept(Dp, R.col, &jb.dropped, c.s);
        k::first_rows((const int32_t*)kcol, n0, jb.kmin, jb.range, first.as<int32_t>(), c.s, kept);
        if (jb.dropped && c.E.cfg.verbose)
          fprintf(stderr, "[svm] %s: the hash table on '%s' drops %llu of %llu inserts\n", c.info.id.c_str(),
                  D.id.c_str(), (unsigned long long)jb.dropped, (unsigned long long)n0);
      }
      k::pack_map(first.as<int32_t>(), jb.range, payload, codes_for_pack.as<int32_t>(), jb.shape.wbits,
                  jb.packed.as<uint32_t>(), c.s);
      // `first` and the codes are released in stream order after pack_map
      finish_map(jb);
    }
    joins.push_back(std::move(jb));
  }
  return joins;
}

// Staged column loads: the tier of each join's key column (KernelShape::join_tier), or
// none.  A column is read in 32-byte sectors (8 rows); a sector of a column loaded
// after tiers that pass an estimated fraction r of rows is read with probability
// 1 - (1 - r)^8.  A join's test costs little when its map or existence bits sit in LDS
// (assign_lds) and about kL2LookupCost columns' worth per row that reaches it when they
// are read from L2 (one scattered load per row: the texture path, not HBM, bounds it).
// Every ordered choice of the dense-map joins that filter (estimated pass fraction
// below 1) is costed: mode 1 puts the chosen joins in one tier, mode 2 gives each its
// own tier in that order; the cheapest plan (plus a small charge per tier for the
// dependent load) is taken, and staging is used only when it costs at most 85% of
// reading every column.  Among equal costs, the most selective joins first.
constexpr double kL2LookupCost = 2.5;

std::vector<int> plan_tiers(const std::vector<JoinBuild>& joins, const KernelShape& ks, int mode) {
  std::vector<size_t> order;
  for (size_t j = 0; j < joins.size(); ++j)
    if (joins[j].shape.kind == JOIN_MAP && joins[j].sel < 1.0) order.push_back(j);
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return joins[x].sel < joins[y].sel; });
  const double ncols = (double)ks.fact_i32.size();
  auto sector = [](double r) { return 1.0 - std::pow(1.0 - r, 8.0); };
  auto in_lds = [&](size_t j) { return joins[j].staged_lds || joins[j].staged_bits; };
  double best = ncols * 0.85;
  std::vector<int> plan;
  // permutations of positions into `order` (the first is `order` itself); each prefix
  // of each permutation is a candidate plan
  std::vector<size_t> perm(order.size());
  for (size_t q = 0; q < perm.size(); ++q) perm[q] = q;
  do {
    for (size_t k = 1; k <= perm.size(); ++k) {
      std::vector<int> tier(joins.size(), -1);
      for (size_t q = 0; q < k; ++q) tier[order[perm[q]]] = mode == 2 ? (int)q : 0;
      // each column at its earliest tier; the rest after the last
      std::map<int, int> ct;
      for (size_t q = 0; q < k; ++q) {
        const int c = joins[order[perm[q]]].shape.lcol, t = tier[order[perm[q]]];
        ct[c] = ct.count(c) ? std::min(ct[c], t) : t;
      }
      const int ntiers = mode == 2 ? (int)k : 1;
      std::vector<double> pass(ntiers + 1, 1.0);  // rows passing the tiers before t
      for (int t = 1; t <= ntiers; ++t) {
        pass[t] = pass[t - 1];
        for (size_t q = 0; q < k; ++q)
          if (tier[order[perm[q]]] == t - 1) pass[t] *= joins[order[perm[q]]].sel;
      }
      double cost = 0.05 * ntiers;
      for (int c : ks.fact_i32) {
        const int t = ct.count(c) ? ct[c] : ntiers;
        cost += t == 0 ? 1.0 : sector(pass[t]);
      }
      for (size_t q = 0; q < k; ++q)
        if (!in_lds(order[perm[q]])) cost += kL2LookupCost * pass[tier[order[perm[q]]]];
      if (cost < best - 1e-9) {
        best = cost;
        plan = tier;
      }
    }
  } while (std::next_permutation(perm.begin(), perm.end()));
  return plan;
}

// LDS layouts, the group table first in both.  Plain kernels (jb.shape.lds, lds_off):
// small dense maps in plan order while they fit.  Staged scans (staged_*): maps and the
// existence bits of deferred maps, the most selective join first (a staged scan tests
// it on every row, the others on the few rows that pass it; among equal estimates,
// plan order).  lds_words[0] / [1]: the words each layout uses.
void assign_lds(std::vector<JoinBuild>& joins, uint64_t agg_bytes, uint64_t lds_words[2]) {
  const uint64_t agg_words = (agg_bytes + 7) / 8 * 2;  // agg table first, 8-byte aligned, in words
  uint64_t used = agg_words;
  for (auto& jb : joins) {
    if (jb.words && !jb.table && !jb.shape.defer && (used + jb.words) * 4 <= kLdsBudget) {
      jb.shape.lds = true;
      jb.lds_off = (uint32_t)used;
      used += jb.words;
    }
  }
  lds_words[0] = used;
  std::vector<size_t> order(joins.size());
  for (size_t j = 0; j < order.size(); ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return joins[x].sel < joins[y].sel; });
  used = agg_words;
  for (size_t j : order) {
    JoinBuild& jb = joins[j];
    if (!jb.words || jb.table) continue;
    const uint64_t need = jb.shape.defer ? k::exist_words(jb.range) : jb.words;
    if ((used + need) * 4 > kLdsBudget) continue;
    (jb.shape.defer ? jb.staged_bits : jb.staged_lds) = true;
    jb.staged_off = (uint32_t)used;
    used += need;
  }
  lds_words[1] = used;
}

void fill_common_args(Ctx& c, SvmArgs& a, const std::vector<JoinBuild>& joins,
                      const std::map<std::pair<int, int>, int>& dcol_slot, const Literals& lits,
                      DevBuf& blob) {
  memset(&a, 0, sizeof(a));
  // source columns by their base address; the kernels scan rows [row0, row0 + rows)
  for (size_t q = 0; q < c.src->cols.size() && q < 24; ++q) a.ptr[q] = c.src->cols[q].dev;
  for (size_t j = 0; j < joins.size(); ++j) {
    if (joins[j].table) {  // probe: the hash table and its size
      a.ptr[SVM_SLOT_MAP + j] = joins[j].table;
      a.iv[4 * j + 1] = (int)(uint32_t)joins[j].slots;
      continue;
    }
    a.ptr[SVM_SLOT_MAP + j] = joins[j].packed.p;
    a.iv[4 * j + 0] = joins[j].kmin;
    a.iv[4 * j + 1] = (int)joins[j].range;
    a.iv[4 * j + 2] = (int)joins[j].words;
    a.iv[4 * j + 3] = (int)joins[j].lds_off;
  }
  for (auto& kv : dcol_slot) {
    const Table& D = *c.refs[kv.first.first - 1];
    if (!D.cols[kv.first.second].dev) fail(SVM_ERR_INVALID, "dimension column is not materialised");
    a.ptr[kv.second] = D.col_ptr(kv.first.second);
  }
  for (size_t q = 0; q < lits.ints.size() && q < SVM_MAX_LIT; ++q) a.pi[q] = lits.ints[q];
  if (!lits.blob.empty()) {
    // zero tail: a string literal may be read up to a projection's width (<= 4096)
    blob = DevBuf(lits.blob.size() + 4096, c.s);
    k::fill_u32(blob.p, blob.bytes / 4, 0, c.s);
    SVM_HIP(hipMemcpyAsync(blob.p, lits.blob.data(), lits.blob.size(), hipMemcpyHostToDevice, c.s));
    a.sblob = blob.as<const char>();
  }
  a.row0 = (long long)c.src->row0;
  a.rows = (long long)c.src->rows;
}

std::string lower_all(Ctx& c, const std::vector<std::string>& exps, Literals& lits,
                      std::set<std::pair<int, int>>& operands, std::vector<std::string>& out) {
  for (auto& e : exps) {
    Lowered l = lower_expr(e, c.schemas, false, &lits);
    operands.insert(l.operands.begin(), l.operands.end());
    out.push_back(l.code);
  }
  return {};
}

void classify_fact_columns(Ctx& c, const std::set<std::pair<int, int>>& ops, KernelShape& ks) {
  std::set<int> i32, other;
  for (auto& o : ops) {
    if (o.first != 0) continue;
    if (c.schemas[0][o.second].first == Int32) i32.insert(o.second);
    else other.insert(o.second);
  }
  for (auto& jb : ks.joins) {
    if (jb.kind == JOIN_MAP || jb.ktype == Int32) i32.insert(jb.lcol);
    else other.insert(jb.lcol);
  }
  for (auto& k : ks.keys)
    if (k.src == KEY_FACT) i32.insert(k.col);
  ks.fact_i32.assign(i32.begin(), i32.end());
  ks.fact_other.assign(other.begin(), other.end());
  for (int col : ks.fact_i32)
    if (col >= 24) fail(SVM_ERR_UNSUPPORTED, "source column index >= 24");
}

std::map<std::pair<int, int>, int> assign_dcols(const std::set<std::pair<int, int>>& ops) {
  std::map<std::pair<int, int>, int> slots;
  int next = SVM_SLOT_DCOL;
  for (auto& o : ops) {
    if (o.first == 0) continue;
    if (next >= SVM_SLOT_CODE) fail(SVM_ERR_UNSUPPORTED, "too many dimension columns in expressions");
    slots[o] = next++;
  }
  return slots;
}

// Workgroups of the persistent scan grid: per CU, `bpc` when given, else as many as LDS
// allows (at most 8), never more than the CU keeps resident at once for this kernel
// (its registers, LDS and workgroup size): a grid larger than that runs its last
// workgroups as a second, partial wave, which streams slower than one full wave.
std::mutex g_occ_mu;
std::map<std::pair<hipFunction_t, uint64_t>, int> g_occ;
int resident_blocks(hipFunction_t fn, int block, uint64_t lds_bytes) {
  std::lock_guard<std::mutex> g(g_occ_mu);
  const auto key = std::make_pair(fn, ((uint64_t)block << 32) | lds_bytes);
  auto it = g_occ.find(key);
  if (it != g_occ.end()) return it->second;
  int n = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, block, (size_t)lds_bytes) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  g_occ[key] = n;
  return n;
}

unsigned scan_grid(Engine& E, hipFunction_t fn, uint64_t vec, int block, int unroll, uint64_t lds_bytes, int bpc) {
  if (bpc <= 0) {
    uint64_t lds_cu = 160 * 1024;
    bpc = lds_bytes ? (int)(lds_cu / (lds_bytes + 512)) : 8;
    if (bpc > 8) bpc = 8;
    if (bpc < 1) bpc = 1;
  }
  const int resident = fn ? resident_blocks(fn, block, lds_bytes) : 0;
  if (resident > 0 && bpc > resident) bpc = resident;
  uint64_t need = (vec + (uint64_t)block * unroll - 1) / ((uint64_t)block * unroll);
  uint64_t cap = (uint64_t)E.num_cus * bpc;
  uint64_t g = need < cap ? need : cap;
  return (unsigned)(g ? g : 1);
}

// ------------------------------------------------------------------ autotuner
// The scan's launch configuration (workgroup size, 128-bit groups per thread per
// iteration, workgroups per CU) is chosen per kernel shape and source size by
// timing the candidates on the first large run of that shape; integer results do
// not depend on it, so the accumulators are simply cleared between trials.
struct TuneChoice {
  int block = 256, unroll = 1, bpc = 0;  // bpc 0: as many as LDS allows (<= 8)
  bool stage = true;                      // small join maps copied into LDS
  bool prefetch = false;                  // software-pipelined column loads
  bool flat = false;                      // branch-free filter + joins for a lane's four rows
  int late = 0;                           // staged column loads (flat form): 1 one tier, 2 chained tiers,
                                          // 3 chained tiers with the passing rows compacted per wave
  bool packed = false;                    // late 3: the tier-0 key column read from its narrow copy
                                          // (unroll counts groups of 8 rows)
};
struct TuneState {
  bool done = false;
  TuneChoice best;
  // time of each candidate; a negative time marks a trial whose group table
  // differed from the first candidate's (never chosen)
  std::vector<std::pair<TuneChoice, float>> trials;
  uint64_t rows = 0;
  std::string pipeline;
};
std::mutex g_tune_mu;
std::map<uint64_t, TuneState> g_tune;
std::map<std::string, std::string> g_last_config;  // pipeline id -> its last scan's configuration
const TuneChoice kTuneCandidates[] = {
    {256, 1, 0, true, false, false},  {256, 2, 0, true, false, false},  {512, 1, 0, true, false, false},
    {512, 1, 0, false, false, false}, {1024, 1, 0, false, false, false}, {512, 1, 0, true, true, false},
    {512, 1, 0, false, true, false},  {512, 2, 0, true, false, false},  {1024, 1, 0, true, false, false},
    {256, 1, 0, true, false, true},   {512, 1, 0, true, false, true},   {512, 2, 0, true, false, true},
    {1024, 1, 0, true, false, true},  {512, 1, 0, false, false, true},
    // two workgroups per CU: a plain read of several columns streams fastest with
    // few waves, a scan with dependent map lookups needs more to hide them
    {512, 1, 2, true, false, false},  {512, 2, 2, true, false, false},
    // the flat form with the next groups' loads in flight during this group's lookups
    {1024, 1, 0, true, true, true},   {512, 1, 0, true, true, true},    {1024, 2, 0, true, false, true},
    // 1,024 threads, one or two workgroups per CU
    {1024, 1, 1, true, false, true},  {1024, 1, 2, true, false, true},
    // staged column loads: the selective joins' key columns first (one tier, or one
    // tier per join), the other columns only for lanes with a row still passing
    {1024, 1, 0, true, true, true, 1},  {1024, 2, 0, true, false, true, 1}, {1024, 1, 0, true, false, true, 2},
    {1024, 1, 0, true, true, true, 2},  {512, 1, 0, true, true, true, 2},   {1024, 1, 2, true, false, true, 2},
    {1024, 2, 0, true, false, true, 2}, {512, 2, 0, true, false, true, 2},  {512, 4, 0, true, false, true, 2},
    {256, 4, 0, true, false, true, 2},
    // the same with the rows that pass tier 0 handed out one per lane
    {512, 2, 0, true, false, true, 3},  {512, 4, 0, true, false, true, 3},  {1024, 4, 0, true, false, true, 3},
    {256, 8, 0, true, false, true, 3},  {512, 8, 0, true, false, true, 3},  {1024, 2, 0, true, false, true, 3},
    {1024, 5, 0, true, false, true, 3}, {512, 6, 0, true, false, true, 3},  {512, 7, 0, true, false, true, 3},
    // the same with the tier-0 key column streamed from its narrow copy (2, 2.25 or 3
    // bytes a row: 8 rows per 128-bit load and one load of their high bits)
    {256, 4, 0, true, false, true, 3, true},  {512, 2, 0, true, false, true, 3, true},
    {512, 4, 0, true, false, true, 3, true},  {1024, 2, 0, true, false, true, 3, true},
    {1024, 4, 0, true, false, true, 3, true}, {512, 1, 0, true, false, true, 3, true},
    {512, 3, 0, true, false, true, 3, true},  {1024, 1, 0, true, false, true, 3, true},
    {1024, 3, 0, true, false, true, 3, true}, {256, 2, 0, true, false, true, 3, true}};
constexpr TuneChoice kCompactDefault = {512, 8, 0, true, false, true, 3};
// Below this many source rows the launch configuration hardly matters and tuning
// would only add compiles to a short query's latency.
constexpr uint64_t kTuneMinRows = 1ull << 26;
constexpr int kNumTuneCandidates = (int)(sizeof(kTuneCandidates) / sizeof(kTuneCandidates[0]));

// SVM_TUNE_MIN_ROWS=<n>: tune shapes from n source rows on (so that the trials and
// their cross-check run on test-sized tables).
uint64_t tune_min_rows() {
  const char* e = getenv("SVM_TUNE_MIN_ROWS");
  if (e && *e) {
    char* end = nullptr;
    const unsigned long long v = strtoull(e, &end, 10);
    if (end && *end == '\0') return v;
  }
  return kTuneMinRows;
}

// SVM_SCAN_VARIANT=<i>: run candidate i of kTuneCandidates, whatever the size, with
// no trials; -1 when unset.
int forced_scan_variant() {
  const char* e = getenv("SVM_SCAN_VARIANT");
  if (!e || !*e) return -1;
  char* end = nullptr;
  const long v = strtol(e, &end, 10);
  if (!end || *end != '\0' || v < 0 || v >= kNumTuneCandidates)
    fail(SVM_ERR_INVALID, std::string("SVM_SCAN_VARIANT must be 0..") + std::to_string(kNumTuneCandidates - 1));
  return (int)v;
}


bool autotune_enabled() {
  const char* e = getenv("SVM_AUTOTUNE");
  return !(e && e[0] == '0');
}

// ------------------------------------------------------------------ host engine
// The source's bulks (its tasks, core/pipeline.cu:161-180), as row ranges relative
// to the source: bulk 0 is the rows of its bulk 0, the rest tuple_per_bulk rows each.
std::vector<std::pair<uint64_t, uint64_t>> source_bulks(const Table& t, uint64_t per_bulk) {
  std::vector<std::pair<uint64_t, uint64_t>> b;
  const uint64_t n = t.rows, b0 = t.bulk0_rows();
  if (!n) return b;
  b.push_back({0, b0 ? b0 : std::min(n, per_bulk)});
  for (uint64_t lo = b.back().second; lo < n; lo += per_bulk) b.push_back({lo, std::min(n, lo + per_bulk)});
  return b;
}

// Which bulks the host engine runs: all of them for a pipeline with isGPU = false;
// with hybrid execution on, bulk i of a GPU pipeline when i % 10 < cpu_ratio
// (core/pipeline.cu:172-180).
std::vector<bool> host_placement(const PipelineInfo& info, size_t nbulks, bool noalloc) {
  const HybridConfig hy = hybrid_config();
  const bool all = !info.isGPU;
  const bool hybrid = !all && hy.enabled && hy.cpu_ratio > 0;
  if ((all || hybrid) && noalloc)
    fail(SVM_ERR_UNSUPPORTED, "svm_pipeline_run_noalloc runs GPU pipelines (isGPU true, hybrid execution off)");
  std::vector<bool> h(nbulks, false);
  for (size_t b = 0; b < nbulks; ++b) h[b] = all || (hybrid && (int)(b % 10) < hy.cpu_ratio);
  return h;
}

// The launch arguments as the host engine reads them: every device buffer the
// generated code reads (the source's referenced columns, the join maps or hash
// tables, dimension columns read by expressions, per-row key codes, literal strings)
// replaced by a host copy (kept alive in `keep`).  Synchronises `c.s` (the maps are
// built there).
SvmArgs host_args(Ctx& c, const KernelShape& ks, const SvmArgs& a, const std::vector<JoinBuild>& joins,
                  const std::vector<KeyBuild>& keys, const Literals& lits,
                  std::vector<std::shared_ptr<const HostBytes>>& keep) {
  SVM_HIP(hipStreamSynchronize(c.s));
  SvmArgs h = a;
  for (auto& p : h.ptr) p = nullptr;
  for (auto& p : h.out) p = nullptr;
  std::vector<int> fcols = ks.fact_i32;
  fcols.insert(fcols.end(), ks.fact_other.begin(), ks.fact_other.end());
  for (int col : fcols) {
    auto m = host_mirror(c.src, col);
    keep.push_back(m);
    h.ptr[col] = m->ptr();
  }
  for (size_t j = 0; j < joins.size(); ++j) {
    std::shared_ptr<HostBytes> b;
    if (joins[j].table) b = host_copy(joins[j].table, joins[j].slots * 4, 16);
    else if (joins[j].packed.p) b = host_copy(joins[j].packed.p, joins[j].packed.bytes, 16);
    if (b) {
      keep.push_back(b);
      h.ptr[SVM_SLOT_MAP + j] = b->ptr();
    }
  }
  for (auto& kv : ks.dcol_slot) {
    const Table& D = *c.refs[kv.first.first - 1];
    const uint64_t size = (uint64_t)D.cols[kv.first.second].size;
    auto b = host_copy(D.col_ptr(kv.first.second), D.rows * size, kColumnPad);
    keep.push_back(b);
    h.ptr[kv.second] = b->ptr();
  }
  for (size_t q = 0; q < keys.size(); ++q)
    if (keys[q].shape.src == KEY_DIMROW && keys[q].codes.p) {
      auto b = host_copy(keys[q].codes.p, keys[q].codes.bytes, 16);
      keep.push_back(b);
      h.ptr[SVM_SLOT_CODE + q] = b->ptr();
    }
  auto blob = std::make_shared<HostBytes>();
  blob->data.reset(new char[lits.blob.size() + 4096]);
  blob->size = lits.blob.size() + 4096;
  memset(blob->data.get(), 0, blob->size);
  if (!lits.blob.empty()) memcpy(blob->data.get(), lits.blob.data(), lits.blob.size());
  keep.push_back(blob);
  h.sblob = blob->ptr();
  return h;
}

// ------------------------------------------------------------------ reduce path
std::unique_ptr<svm_result> run_reduce(Ctx& c, const RunOptions& opt) {
  const Plan& p = c.plan;
  auto res = std::make_unique<svm_result>();
  const int K = (int)p.reduce_key.size(), M = (int)p.metric_exp.size();
  if (M == 0) fail(SVM_ERR_INVALID, "ReduceBy without metrics");
  if (K > 3 || (K > 0 && M > 3) || M > 8)
    fail(SVM_ERR_UNSUPPORTED, "ReduceBy supports 1-3 keys x 1-3 metrics (core/pipeline.cu:380-471) or a scalar");
  const auto& outcols = p.output_info.column_list;
  if ((int)outcols.size() < K + M) fail(SVM_ERR_INVALID, "Sink schema has fewer columns than keys + metrics");
  if (opt.noalloc) {
    if (opt.out_ncols != K + M)
      fail(SVM_ERR_INVALID, "the pipeline's result has " + std::to_string(K + M) + " columns (" + std::to_string(K) +
                                " keys, " + std::to_string(M) + " sums), " + std::to_string(opt.out_ncols) + " given");
    for (int q = 0; q < K + M; ++q)
      if (!opt.out_dev || !opt.out_dev[q]) fail(SVM_ERR_INVALID, "null output column " + std::to_string(q));
    if (opt.out_cap < 1) fail(SVM_ERR_INVALID, "output capacity of 0 rows");
  }
  const auto t_host0 = std::chrono::steady_clock::now();
  Event ev0, ev1, ev2, ev3;
  ev0.record(c.s);

  KernelShape ks;
  ks.block = opt.block;
  ks.unroll = opt.unroll;
  Literals lits;
  std::set<std::pair<int, int>> ops, metric_ops;
  // filters: the reference evaluates them before any join (codegen/codegen.cpp:100-110),
  // so only source columns are defined there.
  std::vector<std::string> fl;
  lower_all(c, p.filter_exp, lits, ops, fl);
  for (auto& o : ops)
    if (o.first != 0) fail(SVM_ERR_UNSUPPORTED, "Filter predicates may only read source columns");
  std::string filt;
  for (size_t q = 0; q < fl.size(); ++q) filt += (q ? "&&" : "") + fl[q];
  ks.filter = fl.empty() ? "true" : filt;
  lower_all(c, p.metric_exp, lits, metric_ops, ks.metrics);
  for (int m = 0; m < M; ++m) {
    const char* t = ctype(outcols[K + m].type);
    if (!t) fail(SVM_ERR_UNSUPPORTED, "metric columns must be integer");
    ks.metric_ctype.push_back(t);
  }
  ops.insert(metric_ops.begin(), metric_ops.end());
  // keys
  std::vector<KeyBuild> keys(K);
  for (int q = 0; q < K; ++q) {
    KeyIndex ki = p.reduce_key[q];
    if (ki.obj < 0 || ki.obj >= (int)c.schemas.size() || ki.col < 0 || ki.col >= (int)c.schemas[ki.obj].size())
      fail(SVM_ERR_INVALID, "reduce key $(" + std::to_string(ki.obj) + "." + std::to_string(ki.col) + ") does not exist");
    if (outcols[q].type != Int32) fail(SVM_ERR_UNSUPPORTED, "reduce key columns must be declared Int32");
    keys[q].shape.obj = ki.obj;
    keys[q].shape.col = ki.col;
    keys[q].shape.src = ki.obj == 0 ? KEY_FACT : KEY_DIMCODE;
    if (ki.obj == 0 && c.schemas[0][ki.col].first != Int32)
      fail(SVM_ERR_UNSUPPORTED, "source-column group keys must be Int32");
  }
  for (int q = 0; q < K; ++q) {
    if (keys[q].shape.obj == 0) continue;
    bool joined = false;
    for (auto& r : p.join_rkey) joined = joined || r.obj == keys[q].shape.obj;
    if (!joined) fail(SVM_ERR_INVALID, "reduce key reads ref object " + std::to_string(keys[q].shape.obj) + " which no JoinWith probes");
  }
  for (auto& o : metric_ops) {
    if (o.first == 0) continue;
    bool joined = false;
    for (auto& r : p.join_rkey) joined = joined || r.obj == o.first;
    if (!joined) fail(SVM_ERR_INVALID, "metric reads ref object " + std::to_string(o.first) + " which no JoinWith probes");
  }
  std::vector<JoinBuild> joins = build_joins(c, metric_ops, keys, /*allow_defer=*/true);
  // probe joins compare against the dimension's key column
  for (auto& jb : joins)
    if (jb.table) metric_ops.insert({jb.shape.robj, jb.shape.rcol});
  const double t_build_host = host_ms(t_host0);
  for (auto& jb : joins) ks.joins.push_back(jb.shape);
  // source-column keys: code = value - min over the column
  for (int q = 0; q < K; ++q) {
    if (keys[q].shape.src != KEY_FACT) continue;
    // range of the source column's origin (cached per resident table)
    auto root = column_root(c.src, keys[q].shape.col);
    KeyEntry ke = c.src->rows ? table_key(root.first, root.second, c.s) : KeyEntry();
    const uint64_t range = c.src->rows ? (uint64_t)((int64_t)ke.kmax - (int64_t)ke.kmin + 1) : 1;
    keys[q].fmin = ke.kmin;
    keys[q].card = range;
    keys[q].n = (uint32_t)std::min<uint64_t>(range, UINT32_MAX);
  }
  // Dense group ids (mixed radix over the codes) while the space is small enough for a
  // group table; beyond that (or a source-column key range above 2^24) the codes are
  // packed into a 64-bit key and the groups found by sort + reduce (AGG_SORT).
  bool wide = false;
  uint64_t G = 1;
  for (int q = 0; q < K; ++q) {
    if (keys[q].shape.src != KEY_FACT) keys[q].card = keys[q].n;
    if (keys[q].card > kMaxFactKeyRange && keys[q].shape.src == KEY_FACT) wide = true;
  }
  for (int q = K - 1; q >= 0 && !wide; --q) {
    keys[q].stride = (uint32_t)G;
    G *= keys[q].card;
    if (G > kMaxGroups) wide = true;
  }
  int key_bits = 0;
  if (wide) {
    G = 0;
    for (int q = K - 1; q >= 0; --q) {
      int b = 0;
      while (b < 32 && (1ull << b) < keys[q].card) b++;
      keys[q].bits = b;
      keys[q].shift = key_bits;
      key_bits += b;
    }
    if (key_bits > 63)
      fail(SVM_ERR_UNSUPPORTED, "group keys need " + std::to_string(key_bits) + " bits of codes (at most 63)");
    if (M > 3) fail(SVM_ERR_UNSUPPORTED, "ReduceBy supports 1-3 metrics (core/pipeline.cu:380-471)");
  }
  for (int q = 0; q < K; ++q) ks.keys.push_back(keys[q].shape);
  ks.dcol_slot = assign_dcols(metric_ops);
  classify_fact_columns(c, ops, ks);
  // aggregation mode and LDS layout
  uint64_t agg_bytes = 0;
  if (K == 0) ks.agg = AGG_SCALAR;
  else if (wide) {
    ks.agg = AGG_SORT;
    ks.hash_bits = kHashBits;
    agg_bytes = (uint64_t)(1 << kHashBits) * (8 + 8 * M);
  } else if (G * M * 8 + (G + 31) / 32 * 4 <= kDenseAggBytes) {
    ks.agg = AGG_DENSE;
    agg_bytes = G * M * 8 + (G + 31) / 32 * 4;
  } else {
    ks.agg = AGG_HASH;
    ks.hash_bits = kHashBits;
    agg_bytes = (uint64_t)(1 << kHashBits) * (4 + 8 * M);
  }
  uint64_t layout_words[2] = {0, 0};
  assign_lds(joins, agg_bytes, layout_words);
  const uint64_t lds_words = layout_words[0];
  for (size_t j = 0; j < joins.size(); ++j) ks.joins[j].lds = joins[j].shape.lds;

  // outputs
  const uint64_t nflag = (G + 31) / 32;
  DevBuf gsum(K ? G * M * 8 + 8 : 8, c.s), gflag(nflag * 4 + 4, c.s), ssum(8 * 8, c.s), entries(16, c.s);
  auto clear_outputs = [&] {
    void* ptrs[4] = {ssum.p, gsum.p, gflag.p, entries.p};
    const uint64_t words[4] = {16, K ? G * M * 2 : 0, K ? nflag : 0, 4};
    k::zero_ranges(ptrs, words, 4, c.s);
  };
  clear_outputs();
  SvmArgs a;
  DevBuf blob;
  fill_common_args(c, a, joins, ks.dcol_slot, lits, blob);
  for (int q = 0; q < K; ++q) {
    a.iv[32 + 4 * q + 0] = keys[q].fmin;
    a.iv[32 + 4 * q + 1] = (int)keys[q].stride;
    a.iv[32 + 4 * q + 2] = (int)keys[q].n;
    a.iv[32 + 4 * q + 3] = keys[q].shift;
    if (keys[q].shape.src == KEY_DIMROW) a.ptr[SVM_SLOT_CODE + q] = keys[q].codes.p;
  }
  a.out[5] = entries.p;
  a.iv[44] = (int)G;
  a.iv[45] = 0;
  a.out[0] = gsum.p;
  a.out[1] = gflag.p;
  a.out[2] = ssum.p;
  const uint64_t rows = c.src->rows;
  const bool run_scan = rows > 0 && (K == 0 || wide || G > 0);
  const uint64_t vec = (rows + 3) / 4;
  TuneChoice choice;
  choice.block = ks.block;
  choice.unroll = ks.unroll;
  choice.bpc = opt.blocks_per_cu;
  // the kernel shape and dynamic LDS of a launch configuration (maps staged or not)
  const uint64_t agg_words = (agg_bytes + 7) / 8 * 2;
  // (a staged scan uses the staged LDS layout; `layout_args` puts its offsets into a
  // copy of the launch arguments)
  // the narrow copy of the tier-0 key column (of the table that holds the column, rows
  // indexed as the scan indexes them), by column
  std::map<int, std::shared_ptr<const PackedKeys>> packs;
  const TablePtr col_owner = c.src->base ? c.src->base : c.src;
  auto shape_for = [&](const TuneChoice& t, KernelShape& k2, uint64_t& words) {
    k2 = ks;
    k2.block = t.block;
    k2.unroll = t.unroll;
    k2.prefetch = t.prefetch;
    k2.flat = t.flat;
    k2.join_tier.clear();
    if (t.late && t.flat) k2.join_tier = plan_tiers(joins, ks, t.late == 1 ? 1 : 2);
    k2.compact = t.late == 3 && !k2.join_tier.empty();
    k2.packed.clear();
    if (t.packed && k2.compact) {
      for (int col : compact_packable(k2)) {
        std::shared_ptr<const PackedKeys> pk;
        if (col_owner->cols[col].dev == c.src->cols[col].dev) {
          auto it = packs.find(col);
          pk = it != packs.end() ? it->second : column_packed_keys(col_owner, col, c.s);
        }
        if (!pk) break;
        packs[col] = pk;
        k2.packed.push_back(col);
        k2.pack_high = pk->high_bits;
      }
      if (k2.packed.empty()) k2.unroll = t.unroll * 2;  // the same rows per lane, as int4 groups
    }
    const bool staged = !k2.join_tier.empty();
    for (size_t j = 0; j < joins.size(); ++j) {
      k2.joins[j].lds = t.stage && (staged ? joins[j].staged_lds : joins[j].shape.lds);
      k2.joins[j].bits_lds = t.stage && staged && joins[j].staged_bits;
    }
    words = t.stage ? layout_words[staged ? 1 : 0] : agg_words;
  };
  auto layout_args = [&](const KernelShape& k2) {
    SvmArgs a2 = a;
    if (!k2.join_tier.empty())
      for (size_t j = 0; j < joins.size(); ++j)
        if (!joins[j].table) a2.iv[4 * j + 3] = (int)joins[j].staged_off;
    if (!k2.packed.empty()) {
      const PackedKeys& pk = *packs.at(k2.packed[0]);
      a2.ptr[SVM_SLOT_PACK] = pk.bytes.p;
      a2.iv[48] = pk.kmin;
    }
    return a2;
  };
  // Untuned default: when the joins are dense maps and staging is estimated to save
  // bytes, the compacting staged form (fastest on every table measured); otherwise the
  // plain form of the requested shape.
  if (run_scan && opt.block == 256 && opt.unroll == 1) {
    bool maps = true;
    for (auto& jb : joins) maps = maps && jb.shape.kind == JOIN_MAP;
    if (maps && !plan_tiers(joins, ks, 2).empty()) choice = kCompactDefault;
  }
  const int forced = forced_scan_variant();
  if (forced >= 0) {
    choice = kTuneCandidates[forced];
  } else if (opt.noalloc) {
    // no trials: the configuration an earlier run of this shape chose, else the default
    if (run_scan && autotune_enabled() && opt.block == 256 && opt.unroll == 1) {
      const uint64_t tkey = scan_source(ks)->hash ^ (uint64_t)(63 - __builtin_clzll(rows));
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune.find(tkey);
      if (it != g_tune.end() && it->second.done) choice = it->second.best;
    }
  } else if (run_scan && !wide && autotune_enabled() && rows >= tune_min_rows() && opt.block == 256 &&
             opt.unroll == 1) {
    const uint64_t tkey = scan_source(ks)->hash ^ (uint64_t)(63 - __builtin_clzll(rows));
    std::unique_lock<std::mutex> lk(g_tune_mu);
    TuneState& st = g_tune[tkey];
    if (!st.done) {
      // Trial runs of every candidate.  Integer results do not depend on the launch
      // shape, so the first candidate's group table is kept and every later trial's
      // is compared with it on the device: a candidate that differs is never chosen.
      const uint64_t sum_words = K ? G * M * 2 : 0, flag_words = K ? nflag : 0;
      DevBuf ref_sum(sum_words * 4 + 8, c.s), ref_flag(flag_words * 4 + 4, c.s), ref_ssum(64, c.s),
          diff(16, c.s);
      st.rows = rows;
      st.pipeline = c.info.id;
      float best = 1e30f;
      for (int ci = 0; ci < kNumTuneCandidates; ++ci) {
        const TuneChoice& cand = kTuneCandidates[ci];
        KernelShape k2;
        uint64_t w2 = 0;
        shape_for(cand, k2, w2);  // (a packed candidate without a narrow tier-0 copy runs as its int4 twin)
        const auto g2src = scan_source(k2, "svm_scan_tune");
        hipFunction_t f2 = jit_function(g2src->text, g2src->hash, "svm_scan_tune");
        unsigned g2 = scan_grid(c.E, f2, vec, k2.block, k2.unroll, w2 * 4, cand.bpc);
        clear_outputs();
        Event ta, tb;
        SVM_HIP(hipEventRecord(ta.e, c.s));
        launch(f2, g2, (unsigned)k2.block, (unsigned)(w2 * 4), c.s, layout_args(k2));
        SVM_HIP(hipEventRecord(tb.e, c.s));
        SVM_HIP(hipEventSynchronize(tb.e));
        float ms = elapsed(ta, tb);
        bool same = true;
        if (ci == 0) {
          SVM_HIP(hipMemcpyAsync(ref_ssum.p, ssum.p, 64, hipMemcpyDeviceToDevice, c.s));
          if (sum_words) SVM_HIP(hipMemcpyAsync(ref_sum.p, gsum.p, sum_words * 4, hipMemcpyDeviceToDevice, c.s));
          if (flag_words) SVM_HIP(hipMemcpyAsync(ref_flag.p, gflag.p, flag_words * 4, hipMemcpyDeviceToDevice, c.s));
        } else {
          same = !k::words_differ(ref_ssum.p, ssum.p, 16, diff.as<uint32_t>(), c.s) &&
                 !k::words_differ(ref_sum.p, gsum.p, sum_words, diff.as<uint32_t>(), c.s) &&
                 !k::words_differ(ref_flag.p, gflag.p, flag_words, diff.as<uint32_t>(), c.s);
        }
        if (!same) {
          fprintf(stderr, "[svm] autotune %s: candidate %d (%d,%d,%d,%d,%d,%d,%d) gave a different group table; "
                          "not chosen\n", c.info.id.c_str(), ci, cand.block, cand.unroll, cand.bpc, (int)cand.stage,
                  (int)cand.prefetch, (int)cand.flat, cand.late);
          ms = -1.0f;
        }
        st.trials.push_back({cand, ms});
        if (same && ms < best) { best = ms; st.best = cand; }
      }
      st.done = true;
      clear_outputs();
      if (c.E.cfg.verbose) {
        fprintf(stderr, "[svm] autotune %s rows=%llu:", c.info.id.c_str(), (unsigned long long)rows);
        for (auto& t : st.trials)
          fprintf(stderr, " (%d,%d,%d,%d,%d,%d,%d)=%.3fms", t.first.block, t.first.unroll, t.first.bpc,
                  (int)t.first.stage, (int)t.first.prefetch, (int)t.first.flat, t.first.late, t.second);
        fprintf(stderr, "\n");
      }
    }
    choice = st.best;
  }
  // SVM_SCAN_FLAT=1 / =0 forces the flat (branch-free join) form on or off
  if (const char* fl = getenv("SVM_SCAN_FLAT")) choice.flat = fl[0] == '1';
  uint64_t run_words = lds_words;
  {
    KernelShape k2;
    shape_for(choice, k2, run_words);
    ks = k2;
    a = layout_args(ks);
  }
  res->tune = std::to_string(ks.block) + "x" + std::to_string(ks.unroll) + (choice.stage ? " lds-maps" : " l2-maps") +
              (choice.prefetch ? " prefetch" : "") + (choice.flat ? " flat" : "");
  {
    // joins that probe the reference's hash table instead of a dense map
    std::string pj;
    for (size_t j = 0; j < joins.size(); ++j)
      if (joins[j].table) pj += (pj.empty() ? "" : ",") + std::to_string(j);
    if (!pj.empty()) res->tune += " probe(" + pj + ")";
  }
  if (!ks.join_tier.empty()) {
    res->tune += choice.late == 3 ? " compacted-tiers(" : choice.late == 2 ? " chained-tiers(" : " staged(";
    for (size_t j = 0; j < ks.join_tier.size(); ++j) res->tune += (j ? "," : "") + std::to_string(ks.join_tier[j]);
    res->tune += ")";
    if (!ks.packed.empty()) {
      res->tune += " packed" + std::to_string(16 + ks.pack_high) + "(";
      for (size_t k = 0; k < ks.packed.size(); ++k) res->tune += (k ? "," : "") + std::to_string(ks.packed[k]);
      res->tune += ")";
    }
  }
  if (run_scan && !opt.noalloc) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_last_config.size() > 4096) g_last_config.clear();
    g_last_config[c.info.id] = res->tune;
  }
  hipFunction_t fn = nullptr;
  if (run_scan) {
    const auto gsrc = scan_source(ks);
    fn = jit_function(gsrc->text, gsrc->hash, "svm_scan");
  }
  const double t_launch_host = host_ms(t_host0);
  ev1.record(c.s);
  // Placement of the source's bulks: the host engine's (all of them for isGPU =
  // false, every bulk i with i % 10 < cpu_ratio under hybrid execution) and runs of
  // consecutive device bulks, one scan launch each.
  const auto bulks = source_bulks(*c.src, c.E.cfg.tuple_per_bulk);
  const std::vector<bool> on_host = host_placement(c.info, bulks.size(), opt.noalloc);
  std::vector<size_t> host_bulks;
  std::vector<std::pair<uint64_t, uint64_t>> dev_runs;
  for (size_t b = 0; b < bulks.size(); ++b) {
    if (on_host[b]) host_bulks.push_back(b);
    else if (!dev_runs.empty() && dev_runs.back().second == bulks[b].first) dev_runs.back().second = bulks[b].second;
    else dev_runs.push_back(bulks[b]);
  }
  const bool dev_scan = run_scan && !dev_runs.empty();
  const bool host_scan = run_scan && !host_bulks.empty();
  // the host engine's arguments and code, ready before the device scan starts so that
  // both run at once
  std::vector<std::shared_ptr<const HostBytes>> keep;
  SvmArgs ha;
  HostTaskFn hfn = nullptr;
  if (host_scan) {
    hfn = host_task_function(gen_host_source(ks));
    ha = host_args(c, ks, a, joins, keys, lits, keep);
  }
  // wide group spaces: the append buffer of (packed key, sums) entries.  Sized first
  // for at most 2^24 passing rows; a scan that needs more (its entry count says so)
  // runs again with the bound: one entry per source row plus one per LDS slot of
  // every workgroup of every launch.
  DevBuf ekeys, evals;
  uint64_t ecap = 0, nentries = 0;
  const unsigned grid = run_scan ? scan_grid(c.E, fn, vec, ks.block, ks.unroll, run_words * 4, choice.bpc) : 0;
  auto size_entries = [&](uint64_t cap) {
    ecap = cap;
    ekeys = DevBuf(cap * 8 + 8, c.s);
    evals = DevBuf(cap * 8 * (uint64_t)M + 8, c.s);
    a.out[3] = ekeys.p;
    a.out[4] = evals.p;
    a.iv[46] = (int)(uint32_t)cap;
    a.iv[47] = (int)(uint32_t)(cap >> 32);
  };
  uint64_t dev_rows = 0;
  for (auto& r : dev_runs) dev_rows += r.second - r.first;
  const uint64_t flush_entries = (uint64_t)dev_runs.size() * grid * (1u << kHashBits);
  const uint64_t entry_bound = dev_rows + flush_entries;
  auto launch_device = [&] {
    for (auto& r : dev_runs) {
      SvmArgs a2 = a;
      a2.row0 = (long long)(c.src->row0 + r.first);
      a2.rows = (long long)(r.second - r.first);
      launch(fn, grid, (unsigned)ks.block, (unsigned)(run_words * 4), c.s, a2);
    }
  };
  if (wide && dev_scan) size_entries(std::min<uint64_t>(entry_bound, (1ull << 24) + flush_entries));
  if (dev_scan) launch_device();
  // the host bulks meanwhile, each a task of the host engine
  const auto t_host_scan = std::chrono::steady_clock::now();
  std::vector<int64_t> hacc(8, 0);
  std::vector<uint64_t> hsum;
  std::vector<uint32_t> hflag;
  std::vector<uint64_t> hkeys;
  std::vector<int64_t> hvals;
  if (host_scan) {
    const HybridConfig hy = hybrid_config();
    // a group table shared by all tasks (atomic adds) when it is large, else one per
    // task, summed into the result as each task ends
    const bool shared_tab = !wide && K > 0 && G * M * 8 > (4ull << 20);
    if (K > 0 && !wide) {
      hsum.assign(G * M, 0);
      hflag.assign(nflag, 0);
    }
    std::mutex merge_mu;
    run_host_tasks(host_bulks.size(), hy.cpu_threads, [&](size_t t) {
      const auto& bk = bulks[host_bulks[t]];
      const uint64_t n = bk.second - bk.first;
      SvmHostOut o;
      memset(&o, 0, sizeof(o));
      std::vector<uint64_t> tsum, tkeys;
      std::vector<uint32_t> tflag;
      std::vector<int64_t> tvals;
      if (K > 0 && !wide) {
        if (shared_tab) {
          o.shared = 1;
          o.gsum = (unsigned long long*)hsum.data();
          o.gflag = hflag.data();
        } else {
          tsum.assign(G * M, 0);
          tflag.assign(nflag, 0);
          o.gsum = (unsigned long long*)tsum.data();
          o.gflag = tflag.data();
        }
      } else if (wide) {
        tkeys.resize(n);
        tvals.resize(n * (uint64_t)M);
        o.ekeys = (unsigned long long*)tkeys.data();
        o.evals = (long long*)tvals.data();
      }
      hfn(&ha, (long long)(c.src->row0 + bk.first), (long long)(c.src->row0 + bk.second), &o);
      std::lock_guard<std::mutex> g(merge_mu);
      for (int m = 0; m < M && K == 0; ++m) hacc[m] += o.acc[m];
      if (K > 0 && !wide && !shared_tab) {
        for (uint64_t i = 0; i < G * M; ++i) hsum[i] += tsum[i];
        for (uint64_t w = 0; w < nflag; ++w) hflag[w] |= tflag[w];
      }
      if (wide) {
        hkeys.insert(hkeys.end(), tkeys.begin(), tkeys.begin() + (long)o.nentries);
        hvals.insert(hvals.end(), tvals.begin(), tvals.begin() + (long)(o.nentries * M));
      }
    });
  }
  const double host_scan_ms = host_scan ? host_ms(t_host_scan) : 0.0;
  if (wide && dev_scan) {
    SVM_HIP(hipMemcpyAsync(&nentries, entries.p, 8, hipMemcpyDeviceToHost, c.s));
    SVM_HIP(hipStreamSynchronize(c.s));
    if (nentries > ecap) {
      size_entries(entry_bound);
      k::fill_u32(entries.p, 2, 0u, c.s);
      launch_device();
      SVM_HIP(hipMemcpyAsync(&nentries, entries.p, 8, hipMemcpyDeviceToHost, c.s));
      SVM_HIP(hipStreamSynchronize(c.s));
      if (nentries > ecap) fail(SVM_ERR_DEVICE, "group entries exceed their bound");
    }
  }
  ev2.record(c.s);
  // the host engine's partial results join the device's
  if (host_scan) {
    if (K == 0) {
      int64_t d[8] = {0};
      SVM_HIP(hipMemcpyAsync(d, ssum.p, 8 * M, hipMemcpyDeviceToHost, c.s));
      SVM_HIP(hipStreamSynchronize(c.s));
      for (int m = 0; m < M; ++m) d[m] = (int64_t)((uint64_t)d[m] + (uint64_t)hacc[m]);
      SVM_HIP(hipMemcpyAsync(ssum.p, d, 8 * M, hipMemcpyHostToDevice, c.s));
      SVM_HIP(hipStreamSynchronize(c.s));
    } else if (!wide) {
      DevBuf tmp(G * M * 8 + nflag * 4 + 8, c.s);
      SVM_HIP(hipMemcpyAsync(tmp.p, hsum.data(), G * M * 8, hipMemcpyHostToDevice, c.s));
      SVM_HIP(hipMemcpyAsync((char*)tmp.p + G * M * 8, hflag.data(), nflag * 4, hipMemcpyHostToDevice, c.s));
      k::add_u64(gsum.p, tmp.p, G * M, c.s);
      k::or_u32(gflag.p, (char*)tmp.p + G * M * 8, nflag, c.s);
      SVM_HIP(hipStreamSynchronize(c.s));
    } else {
      const uint64_t nh = hkeys.size();
      if (nentries + nh > ecap) {
        DevBuf okeys = std::move(ekeys), ovals = std::move(evals);
        size_entries(nentries + nh);
        if (nentries) {
          SVM_HIP(hipMemcpyAsync(ekeys.p, okeys.p, nentries * 8, hipMemcpyDeviceToDevice, c.s));
          SVM_HIP(hipMemcpyAsync(evals.p, ovals.p, nentries * 8 * M, hipMemcpyDeviceToDevice, c.s));
        }
      }
      if (nh) {
        SVM_HIP(hipMemcpyAsync(ekeys.as<uint64_t>() + nentries, hkeys.data(), nh * 8, hipMemcpyHostToDevice, c.s));
        SVM_HIP(hipMemcpyAsync(evals.as<int64_t>() + nentries * M, hvals.data(), nh * 8 * M, hipMemcpyHostToDevice,
                               c.s));
      }
      SVM_HIP(hipStreamSynchronize(c.s));
      nentries += nh;
    }
    res->tune += " host " + std::to_string(host_bulks.size()) + "/" + std::to_string(bulks.size()) + " bulks " +
                 std::to_string(host_scan_ms) + " ms";
  }
  // wide group spaces: sort the entries by key and sum equal keys
  uint64_t wide_rows = 0;
  DevBuf uniq;
  std::vector<DevBuf> wsums;
  int64_t* wsum_ptr[8] = {};
  if (wide) {
    uniq = DevBuf(nentries * 8 + 8, c.s);
    for (int m = 0; m < M; ++m) {
      wsums.emplace_back(nentries * 8 + 8, c.s);
      wsum_ptr[m] = wsums.back().as<int64_t>();
    }
    wide_rows = k::sort_reduce(ekeys.as<uint64_t>(), evals.as<int64_t>(), nentries, M, key_bits,
                               uniq.as<uint64_t>(), wsum_ptr, c.s);
  }
  auto unpack_args = [&](int32_t* const* outs) {
    k::UnpackArgs ua;
    memset(&ua, 0, sizeof(ua));
    ua.K = K;
    for (int q = 0; q < K; ++q) {
      ua.shift[q] = keys[q].shift;
      ua.bits[q] = keys[q].bits;
      ua.dict[q] = keys[q].shape.src == KEY_FACT ? nullptr : keys[q].dict->dict.as<int32_t>();
      ua.fmin[q] = keys[q].fmin;
      ua.out[q] = outs[q];
    }
    return ua;
  };

  // finish: ordered group emission into the output table
  std::vector<std::pair<int, int64_t>> oschema;
  for (int q = 0; q < K; ++q) oschema.push_back({Int32, 4});
  for (int m = 0; m < M; ++m) oschema.push_back({Int64, 8});
  uint64_t out_rows = 1;
  TablePtr out;
  if (opt.noalloc) {
    // into the caller's device buffers; no output object
    if (wide) {
      out_rows = wide_rows;
      if (opt.out_rows) *opt.out_rows = out_rows;
      if (out_rows > opt.out_cap)
        fail(SVM_ERR_INVALID, "the result has " + std::to_string(out_rows) + " groups, the output buffers hold " +
                                  std::to_string(opt.out_cap));
      int32_t* kout[3] = {};
      for (int q = 0; q < K; ++q) kout[q] = (int32_t*)opt.out_dev[q];
      k::unpack_keys(uniq.as<uint64_t>(), out_rows, unpack_args(kout), c.s);
      for (int m = 0; m < M && out_rows; ++m)
        SVM_HIP(hipMemcpyAsync(opt.out_dev[K + m], wsum_ptr[m], out_rows * 8, hipMemcpyDeviceToDevice, c.s));
    } else if (K == 0) {
      for (int m = 0; m < M; ++m)
        SVM_HIP(hipMemcpyAsync(opt.out_dev[m], (char*)ssum.p + 8 * m, 8, hipMemcpyDeviceToDevice, c.s));
    } else {
      uint64_t nb = (nflag + 255) / 256;
      DevBuf blocks(nb * 8 + 8, c.s), offs(nb * 8 + 16, c.s);
      out_rows = G ? k::group_count(gflag.as<uint32_t>(), G, blocks.as<unsigned long long>(),
                                    offs.as<unsigned long long>(), c.s)
                   : 0;
      if (opt.out_rows) *opt.out_rows = out_rows;
      if (out_rows > opt.out_cap)
        fail(SVM_ERR_INVALID, "the result has " + std::to_string(out_rows) + " groups, the output buffers hold " +
                                  std::to_string(opt.out_cap));
      if (out_rows) {
        k::EmitArgs ea;
        memset(&ea, 0, sizeof(ea));
        ea.flags = gflag.as<uint32_t>();
        ea.sums = gsum.as<uint64_t>();
        ea.G = G;
        ea.K = K;
        ea.M = M;
        for (int q = 0; q < K; ++q) {
          ea.n[q] = keys[q].n;
          ea.stride[q] = keys[q].stride;
          ea.fmin[q] = keys[q].fmin;
          ea.dict[q] = keys[q].shape.src == KEY_FACT ? nullptr : keys[q].dict->dict.as<int32_t>();
          ea.keys_out[q] = (int32_t*)opt.out_dev[q];
        }
        for (int m = 0; m < M; ++m) ea.sums_out[m] = (int64_t*)opt.out_dev[K + m];
        k::group_emit(ea, offs.as<unsigned long long>(), c.s);
      }
    }
    if (opt.out_rows) *opt.out_rows = out_rows;
    SVM_HIP(hipStreamSynchronize(c.s));
    res->kind = K == 0 ? 2 : 1;
    res->rows = (int64_t)out_rows;
    res->types = oschema;
    return res;
  }
  if (wide) {
    out_rows = wide_rows;
    out = c.E.store.create(c.info.id, oschema, out_rows, nullptr, c.s);
    int32_t* kout[3] = {};
    for (int q = 0; q < K; ++q) kout[q] = (int32_t*)out->cols[q].dev;
    k::unpack_keys(uniq.as<uint64_t>(), out_rows, unpack_args(kout), c.s);
    for (int m = 0; m < M && out_rows; ++m)
      SVM_HIP(hipMemcpyAsync(out->cols[K + m].dev, wsum_ptr[m], out_rows * 8, hipMemcpyDeviceToDevice, c.s));
  } else if (K == 0) {
    out = c.E.store.create(c.info.id, oschema, 1, nullptr, c.s);
    for (int m = 0; m < M; ++m)
      SVM_HIP(hipMemcpyAsync(out->cols[m].dev, (char*)ssum.p + 8 * m, 8, hipMemcpyDeviceToDevice, c.s));
  } else if (opt.want_host && G && G * row_bytes_of(oschema) <= kOneSyncEmitBytes) {
    // One synchronisation: the group count, the emission into an output sized for
    // every group id and the host copy (the emission also writes each row into
    // host-mapped memory) run back to back; the rows are known when the stream is done.
    uint64_t nb = (nflag + 255) / 256;
    DevBuf blocks(nb * 8 + 8, c.s), offs(nb * 8 + 16, c.s);
    const HostSlots hs = host_slots(c.s);
    hs.host[2] = 0;
    const uint64_t tot = k::group_count_async(gflag.as<uint32_t>(), G, blocks.as<unsigned long long>(),
                                              offs.as<unsigned long long>(), hs.dev + 2, c.s);
    out = c.E.store.create(c.info.id, oschema, G, nullptr, c.s, /*zero_pad=*/false);
    const MappedStaging ms = mapped_staging(c.s, G * row_bytes_of(oschema));
    k::EmitArgs ea;
    memset(&ea, 0, sizeof(ea));
    ea.flags = gflag.as<uint32_t>();
    ea.sums = gsum.as<uint64_t>();
    ea.G = G;
    ea.K = K;
    ea.M = M;
    uint64_t hoff[SVM_MAX_SELECT + 8] = {};
    uint64_t off = 0;
    for (int q = 0; q < K + M; ++q) {
      hoff[q] = off;
      off += G * (uint64_t)oschema[q].second;
    }
    for (int q = 0; q < K; ++q) {
      ea.n[q] = keys[q].n;
      ea.stride[q] = keys[q].stride;
      ea.fmin[q] = keys[q].fmin;
      ea.dict[q] = keys[q].shape.src == KEY_FACT ? nullptr : keys[q].dict->dict.as<int32_t>();
      ea.keys_out[q] = (int32_t*)out->cols[q].dev;
      ea.keys_host[q] = (int32_t*)(ms.dev + hoff[q]);
    }
    for (int m = 0; m < M; ++m) {
      ea.sums_out[m] = (int64_t*)out->cols[K + m].dev;
      ea.sums_host[m] = (int64_t*)(ms.dev + hoff[K + m]);
    }
    ea.total = offs.as<unsigned long long>() + tot;
    k::group_emit(ea, offs.as<unsigned long long>(), c.s);
    out->max_tuple_count = c.E.cfg.tuple_per_bulk;
    ev3.record(c.s);
    SVM_HIP(hipStreamSynchronize(c.s));
    out_rows = hs.host[2];
    if (out_rows > G) fail(SVM_ERR_DEVICE, "group emission: more rows than group ids");
    out->rows = out_rows;
    out->bulk0 = out_rows;
    res->kind = 1;
    res->rows = (int64_t)out_rows;
    res->types = oschema;
    res->cols.clear();
    for (int q = 0; q < K + M; ++q) {
      const char* h = ms.host + hoff[q];
      res->cols.emplace_back(h, h + out_rows * (uint64_t)oschema[q].second);
    }
    res->build_ms = elapsed(ev0, ev1);
    res->scan_ms = elapsed(ev1, ev2);
    res->finish_ms = elapsed(ev2, ev3);
    c.E.store.put(c.info.id, out);
    if (c.E.cfg.verbose)
      fprintf(stderr, "[svm] %s: host %.3f ms (joins built %.3f, launch %.3f); gpu build %.3f scan %.3f finish %.3f; G=%llu agg=%d %s\n",
              c.info.id.c_str(), host_ms(t_host0), t_build_host, t_launch_host, res->build_ms, res->scan_ms,
              res->finish_ms, (unsigned long long)G, ks.agg, res->tune.c_str());
    return res;
  } else {
    uint64_t nb = (nflag + 255) / 256;
    DevBuf blocks(nb * 8 + 8, c.s), offs(nb * 8 + 16, c.s);
    out_rows = G ? k::group_count(gflag.as<uint32_t>(), G, blocks.as<unsigned long long>(),
                                  offs.as<unsigned long long>(), c.s)
                 : 0;
    out = c.E.store.create(c.info.id, oschema, out_rows, nullptr, c.s);
    if (out_rows) {
      k::EmitArgs ea;
      memset(&ea, 0, sizeof(ea));
      ea.flags = gflag.as<uint32_t>();
      ea.sums = gsum.as<uint64_t>();
      ea.G = G;
      ea.K = K;
      ea.M = M;
      for (int q = 0; q < K; ++q) {
        ea.n[q] = keys[q].n;
        ea.stride[q] = keys[q].stride;
        ea.fmin[q] = keys[q].fmin;
        ea.dict[q] = keys[q].shape.src == KEY_FACT ? nullptr : keys[q].dict->dict.as<int32_t>();
        ea.keys_out[q] = (int32_t*)out->cols[q].dev;
      }
      for (int m = 0; m < M; ++m) ea.sums_out[m] = (int64_t*)out->cols[K + m].dev;
      k::group_emit(ea, offs.as<unsigned long long>(), c.s);
    }
  }
  out->max_tuple_count = c.E.cfg.tuple_per_bulk;
  ev3.record(c.s);
  res->kind = K == 0 ? 2 : 1;
  res->rows = (int64_t)out_rows;
  res->types = oschema;
  if (opt.want_host)
    copy_out(*out, oschema, (int)oschema.size(), out_rows, c.s, res->cols);
  else
    SVM_HIP(hipStreamSynchronize(c.s));
  res->build_ms = elapsed(ev0, ev1);
  res->scan_ms = elapsed(ev1, ev2);
  res->finish_ms = elapsed(ev2, ev3);
  c.E.store.put(c.info.id, out);
  if (c.E.cfg.verbose)
    fprintf(stderr, "[svm] %s: host %.3f ms (joins built %.3f, launch %.3f); gpu build %.3f scan %.3f finish %.3f; G=%llu agg=%d %s\n",
            c.info.id.c_str(), host_ms(t_host0), t_build_host, t_launch_host, res->build_ms, res->scan_ms,
            res->finish_ms, (unsigned long long)G, ks.agg, res->tune.c_str());
  return res;
}

// ------------------------------------------------------------------ select path
std::unique_ptr<svm_result> run_select(Ctx& c, const RunOptions& opt) {
  const Plan& p = c.plan;
  auto res = std::make_unique<svm_result>();
  const auto& outcols = p.output_info.column_list;
  const int P = (int)p.select_exp.size();
  if (P > SVM_MAX_SELECT) fail(SVM_ERR_UNSUPPORTED, "at most 10 projected columns");
  if ((int)outcols.size() < P) fail(SVM_ERR_INVALID, "Sink schema has fewer columns than the Map select list");
  const auto t_host0 = std::chrono::steady_clock::now();
  Event ev0, ev1, ev2, ev3;
  ev0.record(c.s);
  KernelShape ks;
  ks.block = 256;
  Literals lits;
  std::set<std::pair<int, int>> ops, proj_ops;
  std::vector<std::string> fl;
  lower_all(c, p.filter_exp, lits, ops, fl);
  for (auto& o : ops)
    if (o.first != 0) fail(SVM_ERR_UNSUPPORTED, "Filter predicates may only read source columns");
  std::string filt;
  for (size_t q = 0; q < fl.size(); ++q) filt += (q ? "&&" : "") + fl[q];
  ks.filter = fl.empty() ? "true" : filt;
  lower_all(c, p.select_exp, lits, proj_ops, ks.proj);
  for (int q = 0; q < P; ++q) {
    int t = outcols[q].type;
    if (t != Byte && t != Int16 && t != Int32 && t != Int64 && t != String)
      fail(SVM_ERR_UNSUPPORTED, "projection type unsupported on the integer path");
    if (outcols[q].val_size <= 0) fail(SVM_ERR_INVALID, "projected column without val_size");
    if (t == String) {
      // StrCopy copies val_size bytes from the operand's row (codegen/codegen.cpp:168);
      // a wider projection reads into the following rows, bounded by the column padding.
      if (outcols[q].val_size > 4096) fail(SVM_ERR_UNSUPPORTED, "string projections wider than 4096 bytes");
      Lowered l = lower_expr(p.select_exp[q], c.schemas, false, nullptr);
      for (auto& o : l.operands)
        if (outcols[q].val_size > c.schemas[o.first][o.second].second + (int64_t)kColumnPad)
          fail(SVM_ERR_UNSUPPORTED, "string projection wider than its source column + padding");
    }
    ks.proj_type.push_back({t, (long long)outcols[q].val_size});
  }
  ops.insert(proj_ops.begin(), proj_ops.end());
  for (auto& o : proj_ops) {
    if (o.first == 0) continue;
    bool joined = false;
    for (auto& r : p.join_rkey) joined = joined || r.obj == o.first;
    if (!joined) fail(SVM_ERR_INVALID, "projection reads ref object " + std::to_string(o.first) + " which no JoinWith probes");
  }
  std::vector<KeyBuild> nokeys;
  std::vector<JoinBuild> joins = build_joins(c, proj_ops, nokeys, /*allow_defer=*/false);
  // probe joins compare against the dimension's key column
  for (auto& jb : joins)
    if (jb.table) proj_ops.insert({jb.shape.robj, jb.shape.rcol});
  for (auto& jb : joins) ks.joins.push_back(jb.shape);
  ks.dcol_slot = assign_dcols(proj_ops);
  classify_fact_columns(c, ops, ks);
  uint64_t layout_words[2] = {0, 0};
  assign_lds(joins, 0, layout_words);
  const uint64_t lds_words = layout_words[0];
  for (size_t j = 0; j < joins.size(); ++j) ks.joins[j].lds = joins[j].shape.lds;

  const uint64_t rows = c.src->rows;
  const uint64_t tile = (uint64_t)ks.block * 4;
  const uint64_t tiles = (rows + tile - 1) / tile;
  if (tiles > 0x7FFFFFFFull) fail(SVM_ERR_UNSUPPORTED, "select source too large");
  // output object: the sink schema; only projected columns are materialised (the
  // reference reserves the trailing column for the join hash table, core/pipeline.cu:94)
  std::vector<std::pair<int, int64_t>> oschema;
  std::vector<bool> mat;
  uint64_t row_bytes = 0;
  for (size_t q = 0; q < outcols.size(); ++q) {
    int64_t vs = outcols[q].val_size > 0 ? outcols[q].val_size : 4;
    oschema.push_back({outcols[q].type, vs});
    mat.push_back((int)q < P);
    if ((int)q < P) row_bytes += (uint64_t)vs;
  }
  // A Map of bare source columns without Filter or JoinWith keeps every row of its
  // source unchanged (in every bulk): the output object shares the source's columns
  // (no copy; it keeps its source alive), holding the bytes the reference's copy holds.
  if (p.filter_exp.empty() && p.join_rkey.empty()) {
    std::vector<int> src_col(P, -1);
    bool bare = P > 0;
    for (int q = 0; q < P && bare; ++q) {
      src_col[q] = bare_source_col(p.select_exp[q], c.schemas, oschema[q].first, oschema[q].second);
      bare = src_col[q] >= 0;
    }
    if (bare) {
      auto out = std::make_shared<Table>();
      out->id = c.info.id;
      out->rows = c.src->rows;
      out->row0 = c.src->row0;
      out->max_tuple_count = c.E.cfg.tuple_per_bulk;
      out->bulk0 = c.src->bulk0_rows();
      out->info.id = c.info.id;
      out->info.storage = 2;  // DeviceMem (com/svm.proto:49-54)
      out->base = c.src->base ? c.src->base : c.src;
      for (size_t q = 0; q < oschema.size(); ++q) {
        Column col;
        col.type = oschema[q].first;
        col.size = oschema[q].second;
        col.owns = false;
        if ((int)q < P) {
          const Column& sc = c.src->cols[src_col[q]];
          col.dev = sc.dev;
          col.bytes = sc.bytes;
          col.origin = c.src;
          col.origin_col = src_col[q];
          col.origin_version = c.src->data_version();
        }
        out->cols.push_back(col);
        ColumnInfo ci;
        ci.type = col.type;
        ci.val_size = col.size;
        out->info.column_list.push_back(ci);
      }
      res->kind = 0;
      res->rows = (int64_t)out->rows;
      for (int q = 0; q < P; ++q) res->types.push_back(oschema[q]);
      if (opt.want_host) copy_out(*out, oschema, P, out->rows, c.s, res->cols);
      c.E.store.put(c.info.id, out);
      if (c.E.cfg.verbose)
        fprintf(stderr, "[svm] %s: select host %.3f ms; every row of '%s' kept, columns shared; rows %llu\n",
                c.info.id.c_str(), host_ms(t_host0), c.src->id.c_str(), (unsigned long long)out->rows);
      return res;
    }
  }
  if (!c.info.isGPU) {
    // The host engine (PipelineInfo.isGPU = false): each bulk of the source is a task
    // that writes its passing rows in row order; the tasks' rows are concatenated in
    // bulk order (sink bulk i holds what source bulk i produced), then copied to HBM.
    (void)host_placement(c.info, 0, opt.noalloc);
    SvmArgs a;
    DevBuf blob;
    fill_common_args(c, a, joins, ks.dcol_slot, lits, blob);
    std::vector<std::shared_ptr<const HostBytes>> keep;
    HostTaskFn hfn = host_task_function(gen_host_source(ks));
    std::vector<KeyBuild> nokeys2;
    const SvmArgs ha = host_args(c, ks, a, joins, nokeys2, lits, keep);
    const auto bulks = source_bulks(*c.src, c.E.cfg.tuple_per_bulk);
    std::vector<std::vector<std::vector<char>>> tbuf(bulks.size());
    std::vector<uint64_t> tn(bulks.size(), 0);
    run_host_tasks(bulks.size(), hybrid_config().cpu_threads, [&](size_t t) {
      const uint64_t n = bulks[t].second - bulks[t].first;
      SvmHostOut o;
      memset(&o, 0, sizeof(o));
      tbuf[t].resize(P);
      for (int q = 0; q < P; ++q) {
        tbuf[t][q].resize(n * (uint64_t)oschema[q].second + 16);
        o.sel[q] = tbuf[t][q].data();
      }
      hfn(&ha, (long long)(c.src->row0 + bulks[t].first), (long long)(c.src->row0 + bulks[t].second), &o);
      tn[t] = o.nsel;
    });
    uint64_t total = 0;
    for (uint64_t n : tn) total += n;
    TablePtr out = c.E.store.create(c.info.id, oschema, total, &mat);
    out->max_tuple_count = c.E.cfg.tuple_per_bulk;
    out->bulk0 = tn.empty() ? 0 : tn[0];
    res->kind = 0;
    res->rows = (int64_t)total;
    for (int q = 0; q < P; ++q) {
      res->types.push_back(oschema[q]);
      const uint64_t size = (uint64_t)oschema[q].second;
      std::vector<uint8_t> host(total * size);
      uint64_t off = 0;
      for (size_t t = 0; t < bulks.size(); ++t) {
        if (tn[t]) memcpy(host.data() + off * size, tbuf[t][q].data(), tn[t] * size);
        off += tn[t];
      }
      if (total) SVM_HIP(hipMemcpy(out->cols[q].dev, host.data(), host.size(), hipMemcpyHostToDevice));
      const int col = bare_source_col(p.select_exp[q], c.schemas, oschema[q].first, oschema[q].second);
      if (col >= 0) {
        out->cols[q].origin = c.src;
        out->cols[q].origin_col = col;
        out->cols[q].origin_version = c.src->data_version();
      }
      if (opt.want_host) res->cols.push_back(std::move(host));
    }
    c.E.store.put(c.info.id, out);
    if (c.E.cfg.verbose)
      fprintf(stderr, "[svm] %s: select on the host engine, %zu bulks, %.3f ms; rows %llu\n", c.info.id.c_str(),
              bulks.size(), host_ms(t_host0), (unsigned long long)total);
    return res;
  }
  // A source of one bulk is compacted into an output sized for every row (trimmed
  // to the passing rows once counted) when that capacity is modest: in one launch up
  // to kSingleSelectTiles tiles, else count + scan + write without a synchronisation
  // between them.  Otherwise the tile counts are scanned first and the output is
  // sized exactly.
  const uint64_t src_bulk0 = c.src->bulk0_rows();
  const bool capacity = tiles && src_bulk0 >= rows && rows * row_bytes <= kSingleSelectBytes;
  const bool single = capacity && tiles <= kSingleSelectTiles;
  SvmArgs a;
  DevBuf blob;
  fill_common_args(c, a, joins, ks.dcol_slot, lits, blob);
  ev1.record(c.s);
  uint64_t total = 0;
  std::shared_ptr<const GeneratedSource> src;
  const HostSlots slots = host_slots(c.s);
  TablePtr out;
  // projected columns into the kernel arguments; a bare source column projected
  // unchanged keeps a subset of that column's values
  auto bind_outputs = [&]() {
    out->max_tuple_count = c.E.cfg.tuple_per_bulk;
    for (int q = 0; q < P; ++q) {
      a.ptr[SVM_SLOT_OUT + q] = out->cols[q].dev;
      const int col = bare_source_col(p.select_exp[q], c.schemas, oschema[q].first, oschema[q].second);
      if (col >= 0) {
        out->cols[q].origin = c.src;
        out->cols[q].origin_col = col;
        out->cols[q].origin_version = c.src->data_version();
      }
    }
  };
  // An output sized for every source row whose count no one reads yet (the result is
  // not returned) is left pending: the count lands in pinned words of its own, and
  // the store waits for it when the table is next looked up, so the host prepares
  // the following pipelines while this one runs.
  PinnedWords pw;
  const bool defer = capacity && !opt.want_host && !opt.noalloc && pinned_words_take(pw);
  std::shared_ptr<PendingRows> pend;
  if (defer) {
    pend = std::make_shared<PendingRows>();
    pend->host = pw.host;
    pend->slot = pw.slot;
    pend->stream = c.s;
  }
  volatile unsigned long long* const count_host = defer ? pw.host : slots.host;
  unsigned long long* const count_dev = defer ? pw.dev : slots.dev;
  auto finish_count = [&]() {
    if (defer) {
      SVM_HIP(hipEventCreateWithFlags(&pend->done, hipEventDisableTiming));
      SVM_HIP(hipEventRecord(pend->done, c.s));
      std::atomic_store(&out->pending, pend);
      return;
    }
    SVM_HIP(hipStreamSynchronize(c.s));
    if (count_host[3]) {
      scan_state_reset(c.s);
      fail(SVM_ERR_DEVICE, "select '" + c.info.id + "': the ordered compaction lost a tile status");
    }
    total = count_host[0];
    // the passing rows (and bulk 0: the source is one bulk)
    out->rows = total;
    out->bulk0 = total;
  };
  if (single) {
    src = select_source(ks);
    hipFunction_t fsingle = jit_function(src->text, src->hash, "svm_sel_single");
    out = c.E.store.create(c.info.id, oschema, rows, &mat, c.s, /*zero_pad=*/false);
    bind_outputs();
    const ScanState st = scan_state(c.s, tiles);
    a.out[5] = count_dev;
    a.out[7] = st.buf;
    a.iv[46] = (int)st.epoch;
    count_host[3] = 0;
    launch(fsingle, (unsigned)tiles, (unsigned)ks.block, (unsigned)(lds_words * 4), c.s, a);
    finish_count();
  } else if (capacity) {
    // count, scan and write back to back into an output sized for every source row,
    // then the row count (the write's last tile ends the output with its padding)
    src = select_source(ks);
    DevBuf counts(tiles * 4 + 4, c.s), offs(tiles * 8 + 16, c.s);
    a.out[6] = offs.p;
    a.out[7] = counts.p;
    out = c.E.store.create(c.info.id, oschema, rows, &mat, c.s, /*zero_pad=*/false);
    bind_outputs();
    count_host[3] = 0;
    launch(jit_function(src->text, src->hash, "svm_sel_count"), (unsigned)tiles, (unsigned)ks.block, (unsigned)(lds_words * 4), c.s, a);
    k::exclusive_scan_u32(counts.as<uint32_t>(), tiles, offs.as<unsigned long long>(),
                          offs.as<unsigned long long>() + tiles, c.s, count_dev + 0);
    launch(jit_function(src->text, src->hash, "svm_sel_write"), (unsigned)tiles, (unsigned)ks.block, (unsigned)(lds_words * 4), c.s, a);
    finish_count();
  } else {
    DevBuf counts(tiles * 4 + 4, c.s), offs(tiles * 8 + 16, c.s);
    a.out[6] = offs.p;
    a.out[7] = counts.p;
    if (tiles) {
      src = select_source(ks);
      hipFunction_t fcount = jit_function(src->text, src->hash, "svm_sel_count");
      launch(fcount, (unsigned)tiles, (unsigned)ks.block, (unsigned)(lds_words * 4), c.s, a);
      k::exclusive_scan_u32(counts.as<uint32_t>(), tiles, offs.as<unsigned long long>(),
                            offs.as<unsigned long long>() + tiles, c.s, slots.dev + 0);
      SVM_HIP(hipStreamSynchronize(c.s));
      total = slots.host[0];
    }
    // Sink bulk i holds what source bulk i produced (core/pipeline.cu:161-162,294-296), so
    // the output's bulk 0 -- all a later join reads of it -- is the passing rows of the
    // source's bulk 0.
    uint64_t out_bulk0 = total;
    if (tiles && src_bulk0 < rows) {
      const uint64_t tiles0 = (src_bulk0 + tile - 1) / tile;
      out_bulk0 = 0;
      if (tiles0) {
        DevBuf counts0(tiles0 * 4 + 4, c.s), offs0(tiles0 * 8 + 16, c.s);
        SvmArgs a0 = a;
        a0.rows = (long long)src_bulk0;
        a0.out[6] = offs0.p;
        a0.out[7] = counts0.p;
        launch(jit_function(src->text, src->hash, "svm_sel_count"), (unsigned)tiles0, (unsigned)ks.block, (unsigned)(lds_words * 4),
               c.s, a0);
        k::exclusive_scan_u32(counts0.as<uint32_t>(), tiles0, offs0.as<unsigned long long>(),
                              offs0.as<unsigned long long>() + tiles0, c.s, slots.dev + 1);
        SVM_HIP(hipStreamSynchronize(c.s));
        out_bulk0 = slots.host[1];
      }
    }
    out = c.E.store.create(c.info.id, oschema, total, &mat, c.s);
    out->bulk0 = out_bulk0;
    bind_outputs();
    if (tiles && total) {
      hipFunction_t fwrite = jit_function(src->text, src->hash, "svm_sel_write");
      launch(fwrite, (unsigned)tiles, (unsigned)ks.block, (unsigned)(lds_words * 4), c.s, a);
    }
  }
  ev2.record(c.s);
  ev3.record(c.s);
  res->kind = 0;
  res->rows = (int64_t)total;
  for (int q = 0; q < P; ++q) res->types.push_back(oschema[q]);
  if (defer) {
    // the result is not returned; its timings would need the launch to finish
    c.E.store.put(c.info.id, out);
    if (c.E.cfg.verbose)
      fprintf(stderr, "[svm] %s: select host %.3f ms; rows pending (capacity %llu)\n", c.info.id.c_str(),
              host_ms(t_host0), (unsigned long long)rows);
    return res;
  }
  if (opt.want_host)
    copy_out(*out, oschema, P, total, c.s, res->cols);
  else
    SVM_HIP(hipStreamSynchronize(c.s));
  res->build_ms = elapsed(ev0, ev1);
  res->scan_ms = elapsed(ev1, ev2);
  res->finish_ms = elapsed(ev2, ev3);
  c.E.store.put(c.info.id, out);
  if (c.E.cfg.verbose)
    fprintf(stderr, "[svm] %s: select host %.3f ms; gpu build %.3f scan %.3f; rows %llu\n", c.info.id.c_str(),
            host_ms(t_host0), res->build_ms, res->scan_ms, (unsigned long long)total);
  return res;
}

}  // namespace

std::string last_scan_config(const std::string& pipeline_id) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_last_config.find(pipeline_id);
  if (it == g_last_config.end()) fail(SVM_ERR_NOT_FOUND, "no scan of pipeline '" + pipeline_id + "' has run");
  return it->second;
}

int scan_tune_candidates() { return kNumTuneCandidates; }

std::string scan_tune_report(bool reset) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  std::string j = "[";
  bool first = true;
  for (auto& kv : g_tune) {
    const TuneState& st = kv.second;
    if (!st.done) continue;
    int rejected = 0;
    std::string tr = "[";
    for (size_t q = 0; q < st.trials.size(); ++q) {
      rejected += st.trials[q].second < 0 ? 1 : 0;
      char b[32];
      snprintf(b, sizeof(b), "%s%.4f", q ? "," : "", st.trials[q].second);
      tr += b;
    }
    tr += "]";
    int best = -1;
    for (int ci = 0; ci < kNumTuneCandidates; ++ci) {
      const TuneChoice& t = kTuneCandidates[ci];
      if (t.block == st.best.block && t.unroll == st.best.unroll && t.bpc == st.best.bpc && t.stage == st.best.stage &&
          t.prefetch == st.best.prefetch && t.flat == st.best.flat && t.late == st.best.late)
        best = ci;
    }
    std::string id;
    for (char ch : st.pipeline) {
      if (ch == '"' || ch == '\\') id += '\\';
      id += (unsigned char)ch < 0x20 ? ' ' : ch;
    }
    j += std::string(first ? "" : ",") + "{\"pipeline\":\"" + id + "\",\"rows\":" + std::to_string(st.rows) +
         ",\"best\":" + std::to_string(best) + ",\"rejected\":" + std::to_string(rejected) + ",\"trials\":" + tr + "}";
    first = false;
  }
  if (reset) g_tune.clear();
  return j + "]";
}

std::unique_ptr<svm_result> run_pipeline(const PipelineInfo& info, const RunOptions& opt) {
  Engine& E = engine();
  Plan plan = decode_plan(info);
  if (plan.has_loader) {
    if (opt.noalloc) fail(SVM_ERR_UNSUPPORTED, "svm_pipeline_run_noalloc runs ReduceBy pipelines only");
    return run_load(info, plan, opt);
  }
  Ctx c{E, plan, info, opt.stream ? opt.stream : E.stream, nullptr, {}, {}};
  // runs on one stream are sequential: the previous run's temporaries on this stream
  // are reused only by this run's work on the same stream (a select that left its row
  // count pending may still be running there)
  arena_reset(c.s);
  resolve(c);
  if (plan.is_reduce && !plan.select_exp.empty())
    fail(SVM_ERR_UNSUPPORTED, "a pipeline with both Map and ReduceBy is not well defined in the reference");
  if (plan.is_reduce) return run_reduce(c, opt);
  if (opt.noalloc) fail(SVM_ERR_UNSUPPORTED, "svm_pipeline_run_noalloc runs ReduceBy pipelines only");
  if (!plan.select_exp.empty()) return run_select(c, opt);
  fail(SVM_ERR_UNSUPPORTED, "pipeline '" + info.id + "' has neither Map nor ReduceBy");
}

}  // namespace svm
