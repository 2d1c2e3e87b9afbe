// capi.cpp — extern "C" entry points of libsvm.so (declared in include/svm.h).
#include <cstring>

#include "exec.h"
#include "host_engine.h"
#include "kernels.h"
#include "stats.h"
#include "svm_internal.h"

namespace svm {
int batch_run(int n, const uint8_t* const* plans, const size_t* lens, svm_result** outs);
int server_start(int workers, uint64_t budget, int max_batch, int mode);
int server_submit(const uint8_t* const* plans, const size_t* lens, int n, uint64_t* ticket);
int server_wait(uint64_t ticket, int timeout_ms, svm_result** out, double* queued_ms, double* total_ms);
int server_stats(uint64_t* queued, uint64_t* running, uint64_t* done, uint64_t* bytes);
int server_stop();
std::string scan_tune_report(bool reset);
std::string last_scan_config(const std::string& pipeline_id);
std::string batch_tune_report(bool reset);
int scan_tune_candidates();
namespace bk {
int scan_variants();
}
}  // namespace svm

using namespace svm;

namespace {
thread_local std::string t_err;

template <typename F>
int guard(F&& f) {
  try {
    f();
    t_err.clear();
    return SVM_OK;
  } catch (const Error& e) {
    t_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    t_err = "host allocation failed";
    return SVM_ERR_NO_MEMORY;
  } catch (const std::exception& e) {
    t_err = e.what();
    return SVM_ERR_INVALID;
  }
}

void need_id(const char* id) {
  if (!id || !*id) fail(SVM_ERR_INVALID, "table id must be a non-empty string");
}
}  // namespace

extern "C" {

const char* svm_last_error(void) { return t_err.c_str(); }
const char* svm_version(void) { return "shadowvm_amd 0.1 (gfx950)"; }

int svm_init(const svm_config* cfg) {
  return guard([&] {
    Config c;
    if (cfg) {
      c.device = cfg->device;
      if (cfg->tuple_per_bulk) c.tuple_per_bulk = cfg->tuple_per_bulk;
      if (cfg->hbm_budget_bytes) c.hbm_budget = cfg->hbm_budget_bytes;
      if (cfg->num_streams > 0) c.num_streams = cfg->num_streams;
      c.verbose = cfg->verbose;
    }
    engine_init(c);
  });
}

int svm_shutdown(void) {
  return guard([&] {
    server_stop();
    engine_shutdown();
  });
}

int svm_set_hybrid(int enable, int cpu_ratio, int cpu_threads) {
  return guard([&] {
    HybridConfig h;
    h.enabled = enable != 0;
    h.cpu_ratio = cpu_ratio;
    h.cpu_threads = cpu_threads;
    set_hybrid_config(h);
  });
}

int svm_table_create(const char* id, int ncols, const int32_t* types, const int64_t* val_sizes,
                     uint64_t rows) {
  return guard([&] {
    need_id(id);
    if (ncols <= 0 || ncols > 24 || !types || !val_sizes)
      fail(SVM_ERR_INVALID, "a table has 1..24 columns (MaxColumnNum, com/cpp/define.hpp:25)");
    std::vector<std::pair<int, int64_t>> s;
    for (int c = 0; c < ncols; ++c) {
      if (types[c] < 0 || types[c] > 8) fail(SVM_ERR_INVALID, "unknown column type");
      s.push_back({types[c], val_sizes[c]});
    }
    Engine& E = engine();
    TablePtr t = E.store.create(id, s, rows);
    t->max_tuple_count = E.cfg.tuple_per_bulk;
    E.store.put(id, t);
  });
}

int svm_table_put_column(const char* id, int col, const void* host_src, uint64_t bytes) {
  return guard([&] {
    need_id(id);
    TablePtr t = engine().store.get(id);
    if (col < 0 || col >= (int)t->cols.size()) fail(SVM_ERR_INVALID, "column index out of range");
    Column& c = t->cols[col];
    uint64_t need = t->rows * (uint64_t)c.size;
    if (bytes != need) fail(SVM_ERR_INVALID, "column needs " + std::to_string(need) + " bytes, got " + std::to_string(bytes));
    if (!c.dev) fail(SVM_ERR_INVALID, "column is not materialised");
    if (!c.owns) fail(SVM_ERR_INVALID, "'" + std::string(id) + "' is a view and read-only");
    if (bytes) {
      // work already queued on any stream (which may read the old bytes) finishes first
      SVM_HIP(hipDeviceSynchronize());
      SVM_HIP(hipMemcpy(c.dev, host_src, bytes, hipMemcpyHostToDevice));
    }
    // facts cached about the old contents are recomputed on next use
    t->version++;
  });
}

int svm_table_get_column(const char* id, int col, void* host_dst, uint64_t bytes) {
  return guard([&] {
    need_id(id);
    TablePtr t = engine().store.get(id);
    if (col < 0 || col >= (int)t->cols.size()) fail(SVM_ERR_INVALID, "column index out of range");
    Column& c = t->cols[col];
    uint64_t have = t->rows * (uint64_t)c.size;
    if (bytes > have) fail(SVM_ERR_INVALID, "column holds only " + std::to_string(have) + " bytes");
    if (!c.dev) fail(SVM_ERR_INVALID, "column is not materialised");
    if (bytes) SVM_HIP(hipMemcpy(host_dst, t->col_ptr(col), bytes, hipMemcpyDeviceToHost));
  });
}

int svm_table_column_device(const char* id, int col, void** dev_ptr) {
  return guard([&] {
    need_id(id);
    TablePtr t = engine().store.get(id);
    if (col < 0 || col >= (int)t->cols.size() || !dev_ptr) fail(SVM_ERR_INVALID, "bad column");
    *dev_ptr = (void*)t->col_ptr(col);
  });
}

int svm_table_info(const char* id, uint64_t* rows, int* ncols) {
  return guard([&] {
    need_id(id);
    TablePtr t = engine().store.get(id);
    if (rows) *rows = t->rows;
    if (ncols) *ncols = (int)t->cols.size();
  });
}

int svm_table_column_type(const char* id, int col, int32_t* type, int64_t* val_size) {
  return guard([&] {
    need_id(id);
    TablePtr t = engine().store.get(id);
    if (col < 0 || col >= (int)t->cols.size()) fail(SVM_ERR_INVALID, "column index out of range");
    if (type) *type = t->cols[col].type;
    if (val_size) *val_size = t->cols[col].size;
  });
}

int svm_table_view(const char* id, const char* base_id, uint64_t row0, uint64_t rows) {
  return guard([&] {
    need_id(id);
    need_id(base_id);
    Engine& E = engine();
    E.store.put(id, make_view(id, E.store.get(base_id), row0, rows));
  });
}

int svm_table_drop(const char* id) {
  return guard([&] {
    need_id(id);
    if (!engine().store.drop(id)) fail(SVM_ERR_NOT_FOUND, std::string("no table '") + id + "'");
    release_dead_derived();
  });
}

int svm_table_exists(const char* id) {
  if (!id || !engine_ready()) return 0;
  return engine().store.find(id) ? 1 : 0;
}

int svm_gen_table(const char* id, int kind, uint64_t rows, uint64_t seed, const svm_gen_params* p) {
  return guard([&] {
    need_id(id);
    gen_device_table(id, kind, rows, seed, p);
  });
}

int svm_gen_host(int kind, uint64_t seed, const svm_gen_params* p, uint64_t row0, uint64_t n,
                 void* const* cols, int threads) {
  return guard([&] {
    if (!cols) fail(SVM_ERR_INVALID, "null column array");
    gen_host(kind, seed, p, row0, n, cols, threads);
  });
}

int svm_gen_schema(int kind, int32_t* types, int64_t* val_sizes) {
  int n = 0;
  int rc = guard([&] {
    auto s = gen_schema(kind);
    n = (int)s.size();
    for (int c = 0; c < n; ++c) {
      if (types) types[c] = s[c].first;
      if (val_sizes) val_sizes[c] = s[c].second;
    }
  });
  return rc == SVM_OK ? n : rc;
}

int svm_pipeline_run(const uint8_t* pb, size_t len, svm_result** out) {
  return guard([&] {
    if (out) *out = nullptr;
    PipelineInfo info = decode_pipeline_info(pb, len);
    Engine& E = engine();
    std::lock_guard<std::mutex> g(E.run_mu);
    RunOptions opt;
    opt.want_host = out != nullptr;
    auto r = run_pipeline(info, opt);
    if (out) *out = r.release();
  });
}

int svm_pipeline_run_noalloc(const uint8_t* pb, size_t len, void* const* out_cols_device, int ncols,
                             uint64_t max_rows, uint64_t* rows) {
  return guard([&] {
    if (rows) *rows = 0;
    PipelineInfo info = decode_pipeline_info(pb, len);
    Engine& E = engine();
    // serialised with every other run on the engine stream: calls never overlap on
    // the scratch region, which keeps the last call's contents
    std::lock_guard<std::mutex> g(E.run_mu);
    NoAllocScope scope(scratch_region(E.cfg.device), kNoAllocScratchBytes);
    RunOptions opt;
    opt.want_host = false;
    opt.noalloc = true;
    opt.out_dev = out_cols_device;
    opt.out_ncols = ncols;
    opt.out_cap = max_rows;
    opt.out_rows = rows;
    (void)run_pipeline(info, opt);
  });
}

int svm_result_kind(const svm_result* r) { return r ? r->kind : -1; }
int64_t svm_result_rows(const svm_result* r) { return r ? r->rows : -1; }
int svm_result_cols(const svm_result* r) { return r ? (int)r->types.size() : -1; }

int svm_result_col_type(const svm_result* r, int col, int32_t* type, int64_t* val_size) {
  return guard([&] {
    if (!r || col < 0 || col >= (int)r->types.size()) fail(SVM_ERR_INVALID, "bad result column");
    if (type) *type = r->types[col].first;
    if (val_size) *val_size = r->types[col].second;
  });
}

const void* svm_result_col(const svm_result* r, int col) {
  if (!r || col < 0 || col >= (int)r->cols.size()) return nullptr;
  return r->cols[col].data();
}

int svm_result_timing(const svm_result* r, double* build_ms, double* scan_ms, double* finish_ms) {
  return guard([&] {
    if (!r) fail(SVM_ERR_INVALID, "null result");
    if (build_ms) *build_ms = r->build_ms;
    if (scan_ms) *scan_ms = r->scan_ms;
    if (finish_ms) *finish_ms = r->finish_ms;
  });
}

int svm_result_describe(const svm_result* r, svm_result_desc* d) {
  return guard([&] {
    if (!r || !d) fail(SVM_ERR_INVALID, "null result or descriptor");
    if (r->types.size() > 24) fail(SVM_ERR_INVALID, "result has more than 24 columns");
    memset(d, 0, sizeof(*d));
    d->kind = r->kind;
    d->ncols = (int32_t)r->types.size();
    d->rows = r->rows;
    for (size_t c = 0; c < r->types.size(); ++c) {
      d->type[c] = r->types[c].first;
      d->val_size[c] = r->types[c].second;
      d->col[c] = c < r->cols.size() ? (const void*)r->cols[c].data() : nullptr;
    }
    d->build_ms = r->build_ms;
    d->scan_ms = r->scan_ms;
    d->finish_ms = r->finish_ms;
  });
}

void svm_result_free(svm_result* r) { delete r; }

int svm_results_describe(int n, svm_result* const* rs, svm_result_desc* d) {
  return guard([&] {
    if (n < 0 || (n && (!rs || !d))) fail(SVM_ERR_INVALID, "bad result arrays");
    for (int i = 0; i < n; ++i) {
      if (!rs[i]) {
        memset(&d[i], 0, sizeof(d[i]));
        d[i].kind = -1;
        continue;
      }
      const int rc = svm_result_describe(rs[i], &d[i]);
      if (rc != SVM_OK) fail(rc, svm_last_error());
    }
  });
}

int svm_results_copy(int n, svm_result* const* rs, void* dst, uint64_t dst_bytes) {
  return guard([&] {
    if (n < 0 || (n && !rs)) fail(SVM_ERR_INVALID, "bad result array");
    uint64_t off = 0;
    for (int i = 0; i < n; ++i) {
      const svm_result* r = rs[i];
      if (!r) continue;
      for (size_t c = 0; c < r->types.size(); ++c) {
        const uint64_t b = (uint64_t)r->rows * (uint64_t)r->types[c].second;
        if (off + b > dst_bytes) fail(SVM_ERR_INVALID, "destination buffer too small");
        if (b && c < r->cols.size()) memcpy((char*)dst + off, r->cols[c].data(), b);
        off = (off + b + 7) & ~7ull;
      }
    }
  });
}

void svm_results_free(int n, svm_result* const* rs) {
  if (!rs) return;
  for (int i = 0; i < n; ++i) delete rs[i];
}

int svm_batch_run(int n, const uint8_t* const* plans, const size_t* lens, svm_result** outs) {
  int rc = SVM_OK;
  int g = guard([&] { rc = batch_run(n, plans, lens, outs); });
  return g != SVM_OK ? g : rc;
}

int svm_server_start(int workers, uint64_t budget_bytes, int max_batch, int mode) {
  return guard([&] { server_start(workers, budget_bytes, max_batch, mode); });
}
int svm_server_submit(const uint8_t* const* plans, const size_t* lens, int n, uint64_t* ticket) {
  return guard([&] { server_submit(plans, lens, n, ticket); });
}
int svm_server_wait(uint64_t ticket, int timeout_ms, svm_result** out) {
  if (out) *out = nullptr;
  int rc = SVM_OK;
  int g = guard([&] { rc = server_wait(ticket, timeout_ms, out, nullptr, nullptr); });
  if (g == SVM_OK && rc == SVM_ERR_BUSY) t_err = "timed out waiting for ticket";
  return g != SVM_OK ? g : rc;
}
int svm_server_wait_timed(uint64_t ticket, int timeout_ms, svm_result** out, double* queued_ms, double* total_ms) {
  if (out) *out = nullptr;
  int rc = SVM_OK;
  int g = guard([&] { rc = server_wait(ticket, timeout_ms, out, queued_ms, total_ms); });
  if (g == SVM_OK && rc == SVM_ERR_BUSY) t_err = "timed out waiting for ticket";
  return g != SVM_OK ? g : rc;
}
int svm_server_stats(uint64_t* queued, uint64_t* running, uint64_t* done, uint64_t* bytes_in_use) {
  return guard([&] { server_stats(queued, running, done, bytes_in_use); });
}
int svm_server_stop(void) {
  return guard([&] { server_stop(); });
}

int svm_debug_copy_column(const char* dst_id, int dst_col, const char* src_id, int src_col) {
  return guard([&] {
    need_id(dst_id);
    need_id(src_id);
    Engine& E = engine();
    TablePtr d = E.store.get(dst_id), s = E.store.get(src_id);
    if (dst_col < 0 || dst_col >= (int)d->cols.size() || src_col < 0 || src_col >= (int)s->cols.size())
      fail(SVM_ERR_INVALID, "column index out of range");
    uint64_t bytes = s->rows * (uint64_t)s->cols[src_col].size;
    if (d->rows * (uint64_t)d->cols[dst_col].size != bytes || !d->cols[dst_col].owns)
      fail(SVM_ERR_INVALID, "destination column must be an owned column of the same byte size");
    k::calib_copy(s->col_ptr(src_col), d->cols[dst_col].dev, bytes & ~15ull, E.stream);
    SVM_HIP(hipStreamSynchronize(E.stream));
  });
}

// Debug: the read-only HBM bound of a scan's columns (svm_read_probe): columns
// cols[0..ncols) of table `id` streamed with non-temporal 128-bit loads at
// `blocks_per_cu` workgroups of `block` threads per CU; best of `reps` launches.
int svm_debug_read_probe(const char* id, int ncols, const int* cols, int block, int blocks_per_cu, int reps,
                         double* best_ms) {
  return guard([&] {
    need_id(id);
    if (!cols || !best_ms || reps < 1) fail(SVM_ERR_INVALID, "bad arguments");
    Engine& E = engine();
    TablePtr t = E.store.get(id);
    std::vector<const void*> ptrs;
    for (int k = 0; k < ncols; ++k) {
      if (cols[k] < 0 || cols[k] >= (int)t->cols.size() || t->cols[cols[k]].type != Int32 || !t->cols[cols[k]].dev)
        fail(SVM_ERR_INVALID, "not a materialised Int32 column");
      ptrs.push_back(t->col_ptr(cols[k]));
    }
    std::lock_guard<std::mutex> g(E.run_mu);
    DevBuf out(64);
    hipEvent_t a, b;
    SVM_HIP(hipEventCreate(&a));
    SVM_HIP(hipEventCreate(&b));
    const unsigned grid = (unsigned)(E.num_cus * std::max(1, blocks_per_cu));
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
      SVM_HIP(hipEventRecord(a, E.stream));
      k::read_probe(ptrs.data(), ncols, t->rows, block, grid, out.p, E.stream);
      SVM_HIP(hipEventRecord(b, E.stream));
      SVM_HIP(hipEventSynchronize(b));
      float ms = 0;
      SVM_HIP(hipEventElapsedTime(&ms, a, b));
      best = std::min(best, ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *best_ms = best;
  });
}

// Debug: inserts the reference's join hash table drops over every row of an Int32
// column inserted in row order (slots = 0: HashRate x the table's bulk capacity).
int svm_debug_hash_drops(const char* id, int col, uint64_t slots, uint64_t* dropped) {
  return guard([&] {
    need_id(id);
    if (!dropped) fail(SVM_ERR_INVALID, "dropped is NULL");
    Engine& E = engine();
    TablePtr t = E.store.get(id);
    if (col < 0 || col >= (int)t->cols.size() || t->cols[col].type != Int32 || !t->cols[col].dev)
      fail(SVM_ERR_INVALID, "not a materialised Int32 column");
    if (!slots) slots = (uint64_t)(uint32_t)(kHashRate * (t->max_tuple_count ? t->max_tuple_count : t->rows));
    *dropped = k::ht_kept((const int32_t*)t->col_ptr(col), t->rows, slots, nullptr, E.stream);
  });
}

// Debug: the autotuners' record as JSON, {"scan": [...], "batch": [...]}: per tuned
// shape the chosen candidate and every trial's time (negative: the trial's group
// tables differed from the first trial's and it was not chosen).  reset != 0 clears
// both caches afterwards, so the next run of each shape tunes again.
int svm_debug_tune_report(char* buf, size_t cap, int reset) {
  return guard([&] {
    if (!buf || cap == 0) fail(SVM_ERR_INVALID, "null output buffer");
    const std::string j = "{\"scan\":" + scan_tune_report(reset != 0) + ",\"batch\":" + batch_tune_report(reset != 0) + "}";
    if (j.size() + 1 > cap) fail(SVM_ERR_INVALID, "output buffer too small (" + std::to_string(j.size() + 1) + " bytes needed)");
    memcpy(buf, j.c_str(), j.size() + 1);
  });
}

int svm_debug_variant_counts(int* scan_candidates, int* batch_variants) {
  return guard([&] {
    if (!scan_candidates || !batch_variants) fail(SVM_ERR_INVALID, "null output");
    *scan_candidates = scan_tune_candidates();
    *batch_variants = bk::scan_variants();
  });
}

int svm_debug_scan_config(const char* pipeline_id, char* buf, size_t cap) {
  return guard([&] {
    if (!pipeline_id || !buf || cap == 0) fail(SVM_ERR_INVALID, "null pipeline id or output buffer");
    const std::string t = last_scan_config(pipeline_id);
    if (t.size() + 1 > cap) fail(SVM_ERR_INVALID, "output buffer too small (" + std::to_string(t.size() + 1) + " bytes needed)");
    memcpy(buf, t.c_str(), t.size() + 1);
  });
}

// Debug: the decoded operator chain of a PipelineInfo as JSON, with every
// expression lowered against the Source operator's declared schema (object 0).
int svm_debug_describe_plan(const uint8_t* pb, size_t len, char* buf, size_t cap) {
  return guard([&] {
    PipelineInfo info = decode_pipeline_info(pb, len);
    Plan p = decode_plan(info);
    auto q = [](const std::string& s) {
      std::string o = "\"";
      for (char ch : s) {
        if (ch == '"' || ch == '\\') o += '\\';
        if ((unsigned char)ch < 0x20) { o += ' '; continue; }
        o += ch;
      }
      return o + "\"";
    };
    Schemas sch(1);
    for (auto& c : p.source_info.column_list) sch[0].push_back({c.type, c.val_size});
    std::string j = "{\"id\":" + q(p.id) + ",\"source_type\":" + q(p.source_type) + ",\"pre\":[";
    for (size_t k = 0; k < info.pre_pipeline_id.size(); ++k) j += (k ? "," : "") + q(info.pre_pipeline_id[k]);
    j += "],\"isGPU\":" + std::string(info.isGPU ? "true" : "false") + ",\"filter\":[";
    Literals lits;
    for (size_t k = 0; k < p.filter_exp.size(); ++k) {
      std::string low;
      try { low = lower_expr(p.filter_exp[k], sch, false, &lits).code; } catch (const Error& e) { low = std::string("!") + e.what(); }
      j += (k ? "," : "") + std::string("[") + q(p.filter_exp[k]) + "," + q(low) + "]";
    }
    j += "],\"select\":[";
    for (size_t k = 0; k < p.select_exp.size(); ++k) j += (k ? "," : "") + q(p.select_exp[k]);
    j += "],\"metric\":[";
    for (size_t k = 0; k < p.metric_exp.size(); ++k) j += (k ? "," : "") + q(p.metric_exp[k]);
    auto ki = [](const std::vector<KeyIndex>& v) {
      std::string o = "[";
      for (size_t k = 0; k < v.size(); ++k) o += (k ? "," : "") + std::string("[") + std::to_string(v[k].obj) + "," + std::to_string(v[k].col) + "]";
      return o + "]";
    };
    j += "],\"reduce_key\":" + ki(p.reduce_key) + ",\"join_lkey\":" + ki(p.join_lkey) +
         ",\"join_rkey\":" + ki(p.join_rkey) + ",\"is_reduce\":" + (p.is_reduce ? "true" : "false") +
         ",\"has_loader\":" + (p.has_loader ? "true" : "false") + ",\"int_literals\":[";
    for (size_t k = 0; k < lits.ints.size(); ++k) j += (k ? "," : "") + std::to_string(lits.ints[k]);
    j += "],\"string_literals\":" + q(lits.blob) + ",\"output\":[";
    for (size_t k = 0; k < p.output_info.column_list.size(); ++k)
      j += (k ? "," : "") + std::string("[") + std::to_string(p.output_info.column_list[k].type) + "," +
           std::to_string(p.output_info.column_list[k].val_size) + "]";
    j += "],\"reencoded_len\":" + std::to_string(encode_pipeline_info(info).size()) + "}";
    if (!buf || cap == 0) fail(SVM_ERR_INVALID, "null output buffer");
    if (j.size() + 1 > cap) fail(SVM_ERR_INVALID, "output buffer too small");
    memcpy(buf, j.c_str(), j.size() + 1);
  });
}

}  // extern "C"
