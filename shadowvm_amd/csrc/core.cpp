// core.cpp — errors, device memory, the HBM column store and the engine singleton.
#include <cstring>

#include "kernels.h"
#include "svm_internal.h"

static_assert(svm::kColumnPad == (uint64_t)svm::k::kPadBytes, "pad kernel clears the column padding");

#include <condition_variable>
#include <exception>
#include <thread>

namespace svm {

void fail(int code, const std::string& msg) { throw Error(code, msg); }

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SVM_ERR_NO_MEMORY : SVM_ERR_DEVICE;
  fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

// Caching allocator: allocations below 1 GiB are rounded up to a power of two and
// returned to a free list on release instead of to the driver (hipFree
// synchronises the device, which would stall every stream in flight); larger ones
// (resident fact columns) go straight to hipMalloc / hipFree.
// Its state is created on first use and never destroyed: caches in other files
// (profiles, dictionaries, ...) are static objects holding device buffers, and
// their destructors release those buffers here at process exit, in whatever order
// the linker gave the files' static objects.
namespace {
struct AllocState {
  std::mutex mu;
  std::map<uint64_t, std::vector<void*>> free_lists;  // bucket size -> free blocks
  std::map<void*, uint64_t> bucket_of;                // live cached block -> bucket size
};
AllocState& alloc_state() {
  static AllocState* const s = new AllocState();
  return *s;
}
constexpr uint64_t kCacheLimit = 1ull << 30;

uint64_t bucket_for(uint64_t b) {
  uint64_t s = 256;
  while (s < b) s <<= 1;
  return s;
}

// The calling thread's no-allocation scope (svm_pipeline_run_noalloc).
struct NoAlloc {
  bool on = false;
  char* base = nullptr;
  uint64_t cap = 0, off = 0;
};
thread_local NoAlloc t_noalloc;

void* raw_alloc(uint64_t bytes) {
  noalloc_refuse("allocating " + std::to_string(bytes) + " bytes of device memory");
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(SVM_ERR_NO_MEMORY, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  }
  return p;
}
}  // namespace

NoAllocScope::NoAllocScope(char* base, uint64_t bytes) {
  if (t_noalloc.on) fail(SVM_ERR_STATE, "no-allocation scopes do not nest");
  t_noalloc.on = true;
  t_noalloc.base = base;
  t_noalloc.cap = bytes;
  t_noalloc.off = 0;
}
NoAllocScope::~NoAllocScope() { t_noalloc = NoAlloc(); }
bool noalloc_active() { return t_noalloc.on; }
void noalloc_refuse(const std::string& what) {
  if (t_noalloc.on)
    fail(SVM_ERR_STATE, "svm_pipeline_run_noalloc would need " + what +
                            "; run the same pipeline shape once with svm_pipeline_run first (it builds the cached "
                            "key ranges, dictionaries, join tables and kernels), or use svm_pipeline_run");
}

void* dev_alloc(uint64_t bytes) {
  noalloc_refuse("a device buffer of " + std::to_string(bytes) + " bytes");
  if (bytes >= kCacheLimit) return raw_alloc(bytes);
  uint64_t b = bucket_for(bytes);
  AllocState& st = alloc_state();
  {
    std::lock_guard<std::mutex> g(st.mu);
    auto& fl = st.free_lists[b];
    if (!fl.empty()) {
      void* p = fl.back();
      fl.pop_back();
      st.bucket_of[p] = b;
      return p;
    }
  }
  void* p = raw_alloc(b);
  std::lock_guard<std::mutex> g(st.mu);
  st.bucket_of[p] = b;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  AllocState& st = alloc_state();
  {
    std::lock_guard<std::mutex> g(st.mu);
    auto it = st.bucket_of.find(p);
    if (it != st.bucket_of.end()) {
      st.free_lists[it->second].push_back(p);
      st.bucket_of.erase(it);
      return;
    }
  }
  (void)hipFree(p);
}

namespace {
struct Arena {
  std::vector<std::pair<char*, uint64_t>> chunks;
  size_t cur = 0;
  uint64_t off = 0;
};
std::mutex g_arena_mu;
std::map<hipStream_t, Arena> g_arenas;
constexpr uint64_t kArenaChunk = 256ull << 20;
}  // namespace

void* arena_alloc(hipStream_t s, uint64_t bytes) {
  bytes = (bytes + 255) & ~255ull;
  if (t_noalloc.on) {
    // the run's temporaries come from the scratch region (a bump pointer; the scope
    // ends with the run)
    if (bytes > t_noalloc.cap - t_noalloc.off)
      fail(SVM_ERR_NO_MEMORY, "the pipeline's temporaries exceed the " + std::to_string(t_noalloc.cap >> 20) +
                                  " MiB scratch region of svm_pipeline_run_noalloc (use svm_pipeline_run)");
    char* p = t_noalloc.base + t_noalloc.off;
    t_noalloc.off += bytes;
    return p;
  }
  std::lock_guard<std::mutex> g(g_arena_mu);
  Arena& a = g_arenas[s];
  while (a.cur < a.chunks.size()) {
    if (a.off + bytes <= a.chunks[a.cur].second) {
      char* p = a.chunks[a.cur].first + a.off;
      a.off += bytes;
      return p;
    }
    a.cur++;
    a.off = 0;
  }
  uint64_t sz = bytes > kArenaChunk ? bytes : kArenaChunk;
  char* p = (char*)raw_alloc(sz);
  a.chunks.push_back({p, sz});
  a.cur = a.chunks.size() - 1;
  a.off = bytes;
  return p;
}

void arena_reset(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_arena_mu);
  Arena& a = g_arenas[s];
  a.cur = 0;
  a.off = 0;
}

namespace {
std::map<hipStream_t, HostSlots> g_host_slots;  // guarded by g_arena_mu
struct ScanBuf {
  unsigned long long* buf = nullptr;
  uint64_t cap = 0;
  bool dirty = true;  // cleared by a kernel on its stream before the next use
};
std::map<hipStream_t, ScanBuf> g_scan_bufs;  // guarded by g_arena_mu
uint32_t g_scan_epoch = 0;                   // guarded by g_arena_mu; unique across buffers
}  // namespace

ScanState scan_state(hipStream_t s, uint64_t tiles) {
  std::lock_guard<std::mutex> g(g_arena_mu);
  ScanBuf& b = g_scan_bufs[s];
  if (b.cap < tiles) {
    noalloc_refuse("a larger ordered-compaction status buffer");
    // an earlier run on the stream may still be compacting (a select whose row count
    // is pending): the old buffer is freed once the stream is idle
    if (b.buf) {
      SVM_HIP(hipStreamSynchronize(s));
      (void)hipFree(b.buf);
    }
    uint64_t cap = 4096;
    while (cap < tiles) cap <<= 1;
    b.buf = (unsigned long long*)raw_alloc((cap + 1) * sizeof(unsigned long long));
    b.cap = cap;
    b.dirty = true;
  }
  if (++g_scan_epoch == 0) {
    // after 2^32 launches the epochs repeat: every buffer starts over from zeros
    for (auto& kv : g_scan_bufs) kv.second.dirty = true;
    g_scan_epoch = 1;
  }
  if (b.dirty) {
    // a kernel on the same stream, so the clearing is ordered before the launch that
    // follows (memory from the allocator may hold another buffer's old statuses;
    // their epochs are older than any epoch handed out from now on)
    k::fill_u32(b.buf, (b.cap + 1) * 2, 0u, s);
    b.dirty = false;
  }
  ScanState st;
  st.buf = b.buf;
  st.epoch = g_scan_epoch;
  return st;
}

namespace {
struct Staging {
  void* p = nullptr;
  uint64_t cap = 0;
};
std::map<hipStream_t, Staging> g_staging;  // guarded by g_arena_mu
}  // namespace

void* host_staging(hipStream_t s, uint64_t bytes) {
  std::lock_guard<std::mutex> g(g_arena_mu);
  Staging& st = g_staging[s];
  if (st.cap < bytes) {
    noalloc_refuse("a larger pinned staging buffer");
    if (st.p) (void)hipHostFree(st.p);
    uint64_t cap = 1 << 20;
    while (cap < bytes) cap <<= 1;
    st.p = nullptr;
    st.cap = 0;
    SVM_HIP(hipHostMalloc(&st.p, cap, hipHostMallocDefault));
    st.cap = cap;
  }
  return st.p;
}

namespace {
struct Mapped {
  void* host = nullptr;
  void* dev = nullptr;
  uint64_t cap = 0;
};
std::map<hipStream_t, Mapped> g_mapped;  // guarded by g_arena_mu
}  // namespace

MappedStaging mapped_staging(hipStream_t s, uint64_t bytes) {
  std::lock_guard<std::mutex> g(g_arena_mu);
  Mapped& st = g_mapped[s];
  if (st.cap < bytes) {
    noalloc_refuse("a larger host-mapped staging buffer");
    if (st.host) (void)hipHostFree(st.host);
    uint64_t cap = 1 << 20;
    while (cap < bytes) cap <<= 1;
    st = Mapped{};
    SVM_HIP(hipHostMalloc(&st.host, cap, hipHostMallocMapped | hipHostMallocCoherent));
    SVM_HIP(hipHostGetDevicePointer(&st.dev, st.host, 0));
    st.cap = cap;
  }
  return MappedStaging{(char*)st.host, (char*)st.dev};
}

void scan_state_reset(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_arena_mu);
  auto it = g_scan_bufs.find(s);
  if (it != g_scan_bufs.end()) it->second.dirty = true;
}

// The pinned word pool: one mapped allocation of kPinnedEntries x 4 words, made on
// first use and freed with the engine's other per-stream state.
// Entries are numbered with the pool's generation, so one taken from a pool since
// freed is never given to its successor.
constexpr int kPinnedEntries = 1024;
static volatile unsigned long long* g_pin_host = nullptr;
static unsigned long long* g_pin_dev = nullptr;
static std::vector<int> g_pin_free;
static int g_pin_gen = 0;

bool pinned_words_take(PinnedWords& w) {
  std::lock_guard<std::mutex> g(g_arena_mu);
  if (noalloc_active()) return false;
  if (!g_pin_host) {
    void* h = nullptr;
    void* d = nullptr;
    if (hipHostMalloc(&h, kPinnedEntries * 4 * sizeof(unsigned long long), hipHostMallocMapped | hipHostMallocCoherent) !=
        hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipHostFree(h);
      return false;
    }
    g_pin_host = (volatile unsigned long long*)h;
    g_pin_dev = (unsigned long long*)d;
    g_pin_gen = (g_pin_gen + 1) & 0x3FFFF;
    g_pin_free.clear();
    for (int i = kPinnedEntries - 1; i >= 0; --i) g_pin_free.push_back(i);
  }
  if (g_pin_free.empty()) return false;
  const int i = g_pin_free.back();
  g_pin_free.pop_back();
  w.slot = g_pin_gen << 12 | i;
  w.host = g_pin_host + 4 * i;
  w.dev = g_pin_dev + 4 * i;
  for (int k = 0; k < 4; ++k) w.host[k] = 0ull;
  return true;
}

void pinned_words_give(int slot) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> g(g_arena_mu);
  if (g_pin_host && (slot >> 12) == g_pin_gen) g_pin_free.push_back(slot & 4095);
}

HostSlots host_slots(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_arena_mu);
  auto it = g_host_slots.find(s);
  if (it != g_host_slots.end()) return it->second;
  noalloc_refuse("the stream's pinned host words");
  void* h = nullptr;
  void* d = nullptr;
  SVM_HIP(hipHostMalloc(&h, kHostSlots * sizeof(unsigned long long), hipHostMallocMapped | hipHostMallocCoherent));
  SVM_HIP(hipHostGetDevicePointer(&d, h, 0));
  memset(h, 0, kHostSlots * sizeof(unsigned long long));
  HostSlots hs;
  hs.host = (volatile unsigned long long*)h;
  hs.dev = (unsigned long long*)d;
  g_host_slots[s] = hs;
  return hs;
}

void arena_release_all() {
  std::lock_guard<std::mutex> g(g_arena_mu);
  for (auto& kv : g_arenas)
    for (auto& c : kv.second.chunks) (void)hipFree(c.first);
  g_arenas.clear();
  for (auto& kv : g_host_slots) (void)hipHostFree((void*)kv.second.host);
  g_host_slots.clear();
  for (auto& kv : g_scan_bufs) (void)hipFree(kv.second.buf);
  g_scan_bufs.clear();
  if (g_pin_host) (void)hipHostFree((void*)g_pin_host);
  g_pin_host = nullptr;
  g_pin_dev = nullptr;
  g_pin_free.clear();
  for (auto& kv : g_staging) (void)hipHostFree(kv.second.p);
  g_staging.clear();
  for (auto& kv : g_mapped) (void)hipHostFree(kv.second.host);
  g_mapped.clear();
}

void* dev_alloc_async(uint64_t bytes, hipStream_t s) {
  noalloc_refuse("allocating " + std::to_string(bytes) + " bytes of device memory");
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
  // a launch may still write the columns
  if (auto p = std::atomic_load(&pending)) p->wait();
  for (auto& c : cols)
    if (c.owns) dev_free(c.dev);
}

void PendingRows::wait() {
  if (done) (void)hipEventSynchronize(done);
}

PendingRows::~PendingRows() {
  wait();
  if (done) (void)hipEventDestroy(done);
  pinned_words_give(slot);
}

void resolve_rows(Table& t) {
  std::shared_ptr<PendingRows> p = std::atomic_load(&t.pending);
  if (!p) return;
  std::lock_guard<std::mutex> g(p->mu);
  if (std::atomic_load(&t.pending) != p) return;  // resolved by another thread meanwhile
  if (p->done) SVM_HIP(hipEventSynchronize(p->done));
  const bool lost = p->host[3] != 0ull;
  const uint64_t n = lost ? 0 : (uint64_t)p->host[0];
  t.rows = n;
  t.bulk0 = n;
  std::atomic_store(&t.pending, std::shared_ptr<PendingRows>());
  if (lost) {
    scan_state_reset(p->stream);
    fail(SVM_ERR_DEVICE, "select '" + t.id + "': the ordered compaction lost a tile status");
  }
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
                       uint64_t rows, const std::vector<bool>* materialise, hipStream_t fill_stream,
                       bool zero_pad) {
  auto t = std::make_shared<Table>();
  t->id = id;
  t->rows = rows;
  t->info.id = id;
  t->info.storage = 2;  // DeviceMem (com/svm.proto:49-54)
  std::vector<char*> pads;
  for (size_t c = 0; c < schema.size(); ++c) {
    Column col;
    col.type = schema[c].first;
    col.size = schema[c].second;
    if (col.size <= 0) fail(SVM_ERR_INVALID, "column " + std::to_string(c) + " has val_size <= 0");
    bool mat = !materialise || (*materialise)[c];
    if (mat) {
      col.bytes = rows * (uint64_t)col.size + kColumnPad;
      col.dev = dev_alloc(col.bytes);
      if (fill_stream) {
        if (zero_pad) pads.push_back((char*)col.dev + rows * (uint64_t)col.size);
      }
      else
        SVM_HIP(hipMemsetAsync(col.dev, 0, col.bytes, nullptr));
    }
    t->cols.push_back(col);
    ColumnInfo ci;
    ci.type = col.type;
    ci.val_size = col.size;
    t->info.column_list.push_back(ci);
  }
  // the padding of every column behind its last row, cleared in one launch
  if (!pads.empty()) k::zero_pads(pads.data(), (int)pads.size(), fill_stream);
  // a table zeroed on the null stream is complete before any stream (none of which
  // waits on the null stream) can fill or read it
  if (!fill_stream) SVM_HIP(hipStreamSynchronize(nullptr));
  return t;
}

void Store::put(const std::string& id, TablePtr t) {
  std::lock_guard<std::mutex> g(mu_);
  tables_[id] = std::move(t);
}

TablePtr Store::get(const std::string& id) {
  TablePtr t;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = tables_.find(id);
    if (it == tables_.end()) fail(SVM_ERR_NOT_FOUND, "no table or pipeline output with id '" + id + "'");
    t = it->second;
  }
  resolve_rows(*t);
  return t;
}

TablePtr Store::find(const std::string& id) {
  TablePtr t;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = tables_.find(id);
    if (it == tables_.end()) return nullptr;
    t = it->second;
  }
  resolve_rows(*t);
  return t;
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

// ----------------------------------------------------------------- host workers
namespace {

class Workers {
 public:
  static constexpr int kThreads = 3;  // plus the caller
  static constexpr size_t kInline = 32;

  void run(size_t n, const std::function<void(size_t)>& fn) {
    std::lock_guard<std::mutex> one(call_mu_);  // one parallel_for at a time
    start();
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      busy_ = kThreads;
      err_ = nullptr;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return busy_ == 0; });
    fn_ = nullptr;
    if (err_) std::rethrow_exception(err_);
  }

 private:
  void start() {
    if (!threads_.empty()) return;
    for (int t = 0; t < kThreads; ++t)
      threads_.emplace_back([this] {
        uint64_t seen = 0;
        for (;;) {
          {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
          }
          work();
          std::lock_guard<std::mutex> g(mu_);
          if (--busy_ == 0) done_.notify_all();
        }
      });
    for (auto& t : threads_) t.detach();  // they live as long as the process
  }
  void work() {
    for (;;) {
      const size_t i = next_.fetch_add(1);
      if (i >= n_) return;
      try {
        (*fn_)(i);
      } catch (...) {
        std::lock_guard<std::mutex> g(mu_);
        if (!err_) err_ = std::current_exception();
      }
    }
  }
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_;
  std::vector<std::thread> threads_;
  const std::function<void(size_t)>* fn_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{0};
  int busy_ = 0;
  uint64_t gen_ = 0;
  std::exception_ptr err_;
};

}  // namespace

void parallel_for(size_t n, const std::function<void(size_t)>& fn) {
  if (n < Workers::kInline) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  static Workers* w = new Workers();  // never destroyed: its threads are detached
  w->run(n, fn);
}

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
  // the device address of this device's no-allocation scratch array
  scratch_init(cfg.device);
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
  arena_release_all();
  (void)hipStreamDestroy(g_engine->stream);
  g_engine.reset();
}

}  // namespace svm
