// svm_internal.h — shared declarations of libsvm's host side.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <functional>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/svm.h"

namespace svm {

// ----------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);
#define SVM_HIP(x) ::svm::hip_check((x), #x)

// ----------------------------------------------------------------- plan format
// Field layout of com/svm.proto:73-144 (ColumnInfo, ObjectInfo, OperatorInfo, PipelineInfo).
enum ValType { Byte = 0, Int16 = 1, Int32 = 2, Int64 = 3, Float16 = 4, Float32 = 5,
               Float64 = 6, String = 7, Tensor = 8 };
enum OpType { OpSource = 0, OpLoader = 1, OpFilter = 2, OpMap = 3, OpReduceBy = 4,
              OpSortBy = 5, OpJoinWith = 6, OpSink = 7 };

struct ColumnInfo {
  std::string id;
  int type = Int32;
  bool isnull = false;
  std::vector<int32_t> dim;
  int64_t val_size = 4;
};
struct ObjectInfo {
  std::string id, host, path;
  std::vector<ColumnInfo> column_list;
  int storage = 0;
};
struct OperatorInfo {
  std::string id;
  int type = 0;
  std::map<std::string, std::string> args;
  // The output schema is decoded for Source, Loader and Sink operators, the only ones
  // whose schema the data plane reads; the others keep its bytes (output_raw).
  ObjectInfo output;
  std::string output_raw;
  bool has_output = false, output_decoded = false;
  int perf = 0;
};
struct PipelineInfo {
  std::string id;
  std::vector<OperatorInfo> workflow;
  std::vector<std::string> pre_pipeline_id;
  bool codegen = false, debug = false, isGPU = false;
};
PipelineInfo decode_pipeline_info(const uint8_t* data, size_t len);
std::string encode_pipeline_info(const PipelineInfo& p);

// ----------------------------------------------------------------- store
struct Table;
using TablePtr = std::shared_ptr<Table>;
struct Column {
  int type = Int32;
  int64_t size = 4;
  void* dev = nullptr;      // nullptr: declared but not materialised (e.g. a hash-table column)
  uint64_t bytes = 0;       // allocation size (data + padding)
  bool owns = true;         // false for the columns of a view
  // A bare-column projection of a select pipeline holds a subset of its source
  // column's values: origin names that column (key ranges / dictionaries of the
  // origin bound this column's).
  std::weak_ptr<Table> origin;
  int origin_col = -1;
  uint64_t origin_version = 0;  // origin's data_version() when the projection was made
};
// A select output whose row count a launch is still computing.  The kernels write the
// count (and the one-launch compaction's lost-status flag) into pinned host words;
// until resolve_rows() has waited for the launch, the table's rows are its capacity.
struct PendingRows {
  hipEvent_t done = nullptr;
  volatile unsigned long long* host = nullptr;  // [0] rows; [3] nonzero: a tile status was lost
  int slot = -1;                                 // its words in the pinned pool
  hipStream_t stream = nullptr;
  std::mutex mu;
  void wait();     // returns once the launch has finished
  ~PendingRows();  // waits, then gives back the event and the words
};

struct Table {
  std::string id;
  uint64_t rows = 0;
  uint64_t row0 = 0;        // views: first row inside the base allocation (column dev = base)
  uint64_t max_tuple_count = 0;
  // Rows of bulk 0, the only bulk a join reads (core/pipeline.cu:91,294-295).  A table
  // cut into bulks of max_tuple_count rows has min(rows, max_tuple_count); a select
  // output has the rows its source's bulk 0 produced (sink bulk i is the output of
  // source bulk i, core/pipeline.cu:161-162,294-296); UINT64_MAX = the default rule.
  uint64_t bulk0 = UINT64_MAX;
  uint64_t bulk0_rows() const {
    if (bulk0 != UINT64_MAX) return bulk0 < rows ? bulk0 : rows;
    return max_tuple_count && max_tuple_count < rows ? max_tuple_count : rows;
  }
  std::vector<Column> cols;
  ObjectInfo info;          // schema as ObjectInfo (ref object info_)
  TablePtr base;            // views keep their base table alive
  // Bumped by every svm_table_put_column on this table: facts cached about its
  // columns (key ranges, dictionaries, kept rows, hash tables, profiles) record the
  // data_version() they were computed at and are recomputed when it has moved.
  std::atomic<uint64_t> version{0};
  // Version of the bytes this table reads: its own, plus its base's for a view or for
  // an object whose columns are shared with its source.
  uint64_t data_version() const { return version.load() + (base ? base->data_version() : 0); }
  // address of row 0 of column c (views offset into the base allocation)
  const char* col_ptr(int c) const {
    return cols[c].dev ? (const char*)cols[c].dev + row0 * (uint64_t)cols[c].size : nullptr;
  }
  // Set while a launch still counts the rows (accessed with std::atomic_load/store);
  // the store resolves it before it hands the table out.
  std::shared_ptr<PendingRows> pending;
  ~Table();
};

// Waits for a pending table's row count and sets rows and bulk 0 (a no-op otherwise).
void resolve_rows(Table& t);

// Pinned host words the device writes and the host reads without a copy: four per
// entry, zeroed when taken.  take() returns false when every entry is in use.
struct PinnedWords {
  volatile unsigned long long* host = nullptr;
  unsigned long long* dev = nullptr;
  int slot = -1;
};
bool pinned_words_take(PinnedWords& w);
void pinned_words_give(int slot);

// Zero-copy view of rows [row0, row0+rows) of `base` (row0 relative to base).
TablePtr make_view(const std::string& id, TablePtr base, uint64_t row0, uint64_t rows);

// Column allocations carry this many zero bytes after the last row, so fixed-width
// copies that read past a row (StrCopy of a wider projection) and 128-bit vector
// loads of a partial last group stay inside the allocation.
constexpr uint64_t kColumnPad = 256;

class Store {
 public:
  // Allocates the columns.  Without a stream every byte is zeroed synchronously
  // (tables filled by the caller); with a stream only the padding after the last
  // row is zeroed, asynchronously, for outputs whose rows a kernel writes in full.
  // zero_pad false: the caller's kernel writes the padding itself (outputs sized to a
  // capacity whose final row count a kernel decides).
  TablePtr create(const std::string& id, const std::vector<std::pair<int, int64_t>>& schema,
                  uint64_t rows, const std::vector<bool>* materialise = nullptr,
                  hipStream_t fill_stream = nullptr, bool zero_pad = true);
  void put(const std::string& id, TablePtr t);
  TablePtr get(const std::string& id);  // throws SVM_ERR_NOT_FOUND
  TablePtr find(const std::string& id);
  bool drop(const std::string& id);
  void clear();
 private:
  std::mutex mu_;
  std::map<std::string, TablePtr> tables_;
};

// ----------------------------------------------------------------- device memory
void* dev_alloc(uint64_t bytes);  // hipMalloc with error mapping (long-lived buffers)
void dev_free(void* p);
// Stream-ordered allocation from the device memory pool (no implicit device
// synchronisation on free; memory is reused by later work on any stream once the
// freeing stream has passed the free).
void* dev_alloc_async(uint64_t bytes, hipStream_t s);
void dev_free_async(void* p, hipStream_t s);
// Per-stream bump arena for the temporaries of one pipeline run: allocation is a
// pointer bump; everything is released at once by arena_reset(s), which a run calls
// when it starts (the previous run on that stream has synchronised).
void* arena_alloc(hipStream_t s, uint64_t bytes);
void arena_reset(hipStream_t s);
void arena_release_all();
// Pinned, coherent host words that kernels write directly: counts and totals the
// host reads after synchronising the stream, with no copy dispatch.  Each stream has
// kHostSlots of them; a run owns its stream's slots until it has synchronised.
// Slot owners: 0 = select output rows, 1 = select bulk-0 rows, 2 = group count,
// 3 = one-launch select failure (a tile found no consistent status in time).
constexpr int kHostSlots = 16;
struct HostSlots {
  volatile unsigned long long* host = nullptr;
  unsigned long long* dev = nullptr;
};
HostSlots host_slots(hipStream_t s);
// A stream's pinned host staging buffer of at least `bytes` bytes (grown on demand;
// a run owns it until it has synchronised).  Used to copy results out with one
// synchronisation for all columns.
void* host_staging(hipStream_t s, uint64_t bytes);
// A stream's host-mapped, coherent pinned buffer of at least `bytes` bytes that
// kernels write directly (host and device addresses of the same memory; grown on
// demand; a run owns it until it has synchronised).
struct MappedStaging {
  char* host = nullptr;
  char* dev = nullptr;
};
MappedStaging mapped_staging(hipStream_t s, uint64_t bytes);
// Per-stream state of the one-launch ordered compaction (a tile ticket, which each
// launch leaves at 0, and one status word per tile, tagged with a launch epoch so
// the statuses are never cleared between launches): buf[0] is the ticket,
// buf[1..tiles] the statuses.  Each call returns
// an epoch no earlier call returned (for any stream); the buffer holds at least
// `tiles` statuses and is cleared in stream order when new.
struct ScanState {
  unsigned long long* buf = nullptr;
  uint32_t epoch = 0;
};
ScanState scan_state(hipStream_t s, uint64_t tiles);
// Clears the stream's state before its next use (after a launch that failed).
void scan_state_reset(hipStream_t s);
// Temporary device buffer.  With a stream it lives in that stream's arena until
// the next arena_reset; without one it is a plain (cached) device allocation.
struct DevBuf {
  void* p = nullptr;
  uint64_t bytes = 0;
  hipStream_t s = nullptr;
  bool async = false;
  DevBuf() = default;
  explicit DevBuf(uint64_t b) : p(b ? dev_alloc(b) : nullptr), bytes(b) {}
  DevBuf(uint64_t b, hipStream_t st) : p(b ? arena_alloc(st, b) : nullptr), bytes(b), s(st), async(true) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), s(o.s), async(o.async) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; bytes = o.bytes; s = o.s; async = o.async;
      o.p = nullptr; o.bytes = 0;
    }
    return *this;
  }
  void release() {
    if (!p) return;
    if (!async) dev_free(p);  // arena memory is reclaimed by arena_reset
    p = nullptr;
  }
  ~DevBuf() { release(); }
  template <typename T> T* as() const { return (T*)p; }
};

// ----------------------------------------------------------------- no-allocation runs
// svm_pipeline_run_noalloc runs a fact pipeline without allocating: while a
// NoAllocScope is alive on the calling thread, arena_alloc hands out memory of the
// device scratch region (scratch.hip) instead of the stream arenas, and everything
// that would allocate device or pinned host memory, create an event or load a code
// object (dev_alloc, host staging, a JIT cache miss, ...) fails with SVM_ERR_STATE
// (the pipeline's cached state -- key ranges, dictionaries, compiled kernels -- is
// built by one svm_pipeline_run of the same shape beforehand).
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
void scratch_init(int device);     // once per device, with the device current
char* scratch_region(int device);  // device address of that device's scratch array
struct NoAllocScope {
  NoAllocScope(char* base, uint64_t bytes);
  ~NoAllocScope();
  NoAllocScope(const NoAllocScope&) = delete;
  NoAllocScope& operator=(const NoAllocScope&) = delete;
};
bool noalloc_active();
// Fails with SVM_ERR_STATE, naming `what`, when a NoAllocScope is active.
void noalloc_refuse(const std::string& what);

// ----------------------------------------------------------------- engine
struct Config {
  int device = 0;
  uint64_t tuple_per_bulk = 3000000;
  uint64_t hbm_budget = 32ull << 30;
  int num_streams = 8;
  int verbose = 0;
};

struct Engine {
  Config cfg;
  Store store;
  int num_cus = 256;
  int max_lds = 65536;
  hipStream_t stream = nullptr;  // default execution stream of svm_pipeline_run
  std::mutex run_mu;             // serialises plain pipeline runs on `stream`
};
Engine& engine();  // throws SVM_ERR_STATE if not initialised
bool engine_ready();

// Runs fn(0..n-1) on the calling thread and a few persistent host workers; returns
// when all are done and rethrows the first exception any of them raised.  Small n
// runs inline.
void parallel_for(size_t n, const std::function<void(size_t)>& fn);

}  // namespace svm
