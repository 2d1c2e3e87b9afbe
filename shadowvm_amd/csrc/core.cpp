// core.cpp — errors, device memory, the HBM column store and the engine singleton.
#include <cstring>

#include "kernels.h"
#include "svm_internal.h"

namespace svm {

void fail(int code, const std::string& msg) { throw Error(code, msg); }

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SVM_ERR_NO_MEMORY : SVM_ERR_DEVICE;
  fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

static std::atomic<uint64_t> g_live_bytes{0};

void* dev_alloc(uint64_t bytes) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(SVM_ERR_NO_MEMORY, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  }
  g_live_bytes += bytes;
  return p;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

void* dev_alloc_async(uint64_t bytes, hipStream_t s) {
  void* p = nullptr;
  hipError_t e = hipMallocAsync(&p, bytes ? bytes : 1, s);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(SVM_ERR_NO_MEMORY, "hipMallocAsync(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  }
  return p;
}

void dev_free_async(void* p, hipStream_t s) {
  if (p) (void)hipFreeAsync(p, s);
}

Table::~Table() {
  for (auto& c : cols)
    if (c.owns) dev_free(c.dev);
}

TablePtr make_view(const std::string& id, TablePtr base, uint64_t row0, uint64_t rows) {
  if (row0 + rows > base->rows) fail(SVM_ERR_INVALID, "view rows exceed table '" + base->id + "'");
  auto v = std::make_shared<Table>();
  v->id = id;
  v->rows = rows;
  v->row0 = base->row0 + row0;
  v->max_tuple_count = base->max_tuple_count;
  v->cols = base->cols;
  for (auto& c : v->cols) c.owns = false;
  v->info = base->info;
  v->info.id = id;
  v->base = base->base ? base->base : base;
  return v;
}

TablePtr Store::create(const std::string& id, const std::vector<std::pair<int, int64_t>>& schema,
                       uint64_t rows, const std::vector<bool>* materialise, hipStream_t fill_stream) {
  auto t = std::make_shared<Table>();
  t->id = id;
  t->rows = rows;
  t->info.id = id;
  t->info.storage = 2;  // DeviceMem (com/svm.proto:49-54)
  for (size_t c = 0; c < schema.size(); ++c) {
    Column col;
    col.type = schema[c].first;
    col.size = schema[c].second;
    if (col.size <= 0) fail(SVM_ERR_INVALID, "column " + std::to_string(c) + " has val_size <= 0");
    bool mat = !materialise || (*materialise)[c];
    if (mat) {
      col.bytes = rows * (uint64_t)col.size + kColumnPad;
      col.dev = dev_alloc(col.bytes);
      if (fill_stream)
        SVM_HIP(hipMemsetAsync((char*)col.dev + rows * (uint64_t)col.size, 0, kColumnPad, fill_stream));
      else
        SVM_HIP(hipMemset(col.dev, 0, col.bytes));
    }
    t->cols.push_back(col);
    ColumnInfo ci;
    ci.type = col.type;
    ci.val_size = col.size;
    t->info.column_list.push_back(ci);
  }
  return t;
}

void Store::put(const std::string& id, TablePtr t) {
  std::lock_guard<std::mutex> g(mu_);
  tables_[id] = std::move(t);
}

TablePtr Store::get(const std::string& id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = tables_.find(id);
  if (it == tables_.end()) fail(SVM_ERR_NOT_FOUND, "no table or pipeline output with id '" + id + "'");
  return it->second;
}

TablePtr Store::find(const std::string& id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = tables_.find(id);
  return it == tables_.end() ? nullptr : it->second;
}

bool Store::drop(const std::string& id) {
  std::lock_guard<std::mutex> g(mu_);
  return tables_.erase(id) > 0;
}

void Store::clear() {
  std::lock_guard<std::mutex> g(mu_);
  tables_.clear();
}

static std::unique_ptr<Engine> g_engine;
static std::mutex g_engine_mu;

Engine& engine() {
  if (!g_engine) fail(SVM_ERR_STATE, "svm_init has not been called");
  return *g_engine;
}
bool engine_ready() { return (bool)g_engine; }

void engine_init(const Config& cfg) {
  std::lock_guard<std::mutex> g(g_engine_mu);
  if (g_engine) fail(SVM_ERR_STATE, "svm_init called twice without svm_shutdown");
  int n = 0;
  SVM_HIP(hipGetDeviceCount(&n));
  if (cfg.device < 0 || cfg.device >= n)
    fail(SVM_ERR_DEVICE, "device " + std::to_string(cfg.device) + " not present (" + std::to_string(n) + " visible)");
  SVM_HIP(hipSetDevice(cfg.device));
  auto e = std::make_unique<Engine>();
  e->cfg = cfg;
  hipDeviceProp_t prop;
  SVM_HIP(hipGetDeviceProperties(&prop, cfg.device));
  e->num_cus = prop.multiProcessorCount;
  e->max_lds = (int)prop.sharedMemPerBlock;
  SVM_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  // keep freed stream-ordered memory in the pool instead of returning it to the
  // driver at every synchronisation
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, cfg.device) == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  (void)hipGetLastError();
  g_engine = std::move(e);
}

void engine_shutdown() {
  std::lock_guard<std::mutex> g(g_engine_mu);
  if (!g_engine) return;
  (void)hipStreamSynchronize(g_engine->stream);
  g_engine->store.clear();
  (void)hipStreamDestroy(g_engine->stream);
  g_engine.reset();
}

}  // namespace svm
