/*
 * svm.h — C ABI of the MI355X ShadowVM data plane (libsvm.so).
 *
 * The reference executes one plan through the gRPC method
 *   ShadowVMAPI.RunPipelineCMD(PipelineInfo) -> Ret      (com/svm.proto:157-159,
 *                                                          service/service.cu:25-63)
 * which resolves pre_pipeline_id[0] (source) and [1..] (ref objects) in an
 * id-keyed object cache, runs PipelineTest::Run (test/test.hpp:122-146 ->
 * Pipeline::Pipeline/Open/Run, core/pipeline.cu:11-254) and caches the sink
 * object under PipelineInfo.id.  svm_pipeline_run() is that call; the table
 * functions replace the reference's object store (object/store.cpp:13-91,
 * object/object.cpp:55-294) with a device-resident column store in HBM.
 *
 * Conventions: every function returns 0 (SVM_OK) or a negative svm_status;
 * svm_last_error() gives this thread's message for the last failure.  Calls are
 * thread-safe.  Pointers are plain host pointers unless a name says _device.
 */
#ifndef SVM_H_
#define SVM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum svm_status {
  SVM_OK = 0,
  SVM_ERR_INVALID = -1,      /* malformed argument, plan or expression */
  SVM_ERR_NOT_FOUND = -2,    /* unknown table / pre_pipeline_id (reference: CANCELLED) */
  SVM_ERR_UNSUPPORTED = -3,  /* plan shape outside the device path */
  SVM_ERR_DEVICE = -4,       /* HIP / hipRTC failure */
  SVM_ERR_NO_MEMORY = -5,    /* HBM budget or allocation failure */
  SVM_ERR_JIT = -6,          /* generated kernel failed to compile */
  SVM_ERR_BUSY = -7,         /* admission refused (serving) */
  SVM_ERR_STATE = -8         /* library not initialised / already initialised */
} svm_status;

/* Column value types, numbered as ValType in com/svm.proto:62-72. */
enum { SVM_BYTE = 0, SVM_INT16 = 1, SVM_INT32 = 2, SVM_INT64 = 3, SVM_FLOAT16 = 4,
       SVM_FLOAT32 = 5, SVM_FLOAT64 = 6, SVM_STRING = 7, SVM_TENSOR = 8 };

typedef struct svm_config {
  int device;                 /* HIP device ordinal */
  uint64_t tuple_per_bulk;    /* bulk capacity; dimension max_tuple_count (ref Context::tuple_per_bulk_, com/cpp/com.cu:13) */
  uint64_t hbm_budget_bytes;  /* cap for hash maps + intermediates of admitted queries (0 = 32 GiB) */
  int num_streams;            /* stream pool size (ref ResMgr::stream_num_ = 8, core/resource.cu:27) */
  int verbose;                /* 1: print per-pipeline timing lines to stderr */
} svm_config;

/* ---- lifecycle (ref ResMgr::Init, core/resource.cu:35-47) ---- */
int svm_init(const svm_config* cfg);  /* cfg may be NULL for defaults */
int svm_shutdown(void);
const char* svm_last_error(void);
const char* svm_version(void);

/* ---- host engine and hybrid execution (ref ResMgr::enable_hybrid_exec_, CPU_ratio_,
 * CPU_num_, core/resource.cu:14-21; Pipeline::Run, core/pipeline.cu:139-230) ----
 * A pipeline whose PipelineInfo.isGPU is false runs its per-row transform on host
 * cores: code generated from the same lowered plan as the kernels, compiled by the
 * host C++ compiler at run time, one task per bulk of the source.  With hybrid
 * execution enabled, bulk i of a GPU ReduceBy pipeline runs on the host when
 * i % 10 < cpu_ratio and on the device otherwise (both at once); the host's partial
 * group tables are merged into the device's.  At most cpu_threads host tasks run at
 * once.  Process-wide, like the reference's ResMgr statics (defaults: off, 5, 16). */
int svm_set_hybrid(int enable, int cpu_ratio, int cpu_threads);

/* ---- device column store (ref ObjectStore / RamObject, object/store.cpp:13-91) ---- */
/* Create an empty table of `rows` rows; types/val_sizes follow ColumnInfo (com/svm.proto:73-79). */
int svm_table_create(const char* id, int ncols, const int32_t* types, const int64_t* val_sizes,
                     uint64_t rows);
/* Copy one column from host memory (rows * val_size bytes) into HBM. */
int svm_table_put_column(const char* id, int col, const void* host_src, uint64_t bytes);
/* Copy one column out of HBM into host memory. */
int svm_table_get_column(const char* id, int col, void* host_dst, uint64_t bytes);
/* Device address of a column (valid until the table is dropped). */
int svm_table_column_device(const char* id, int col, void** dev_ptr);
int svm_table_info(const char* id, uint64_t* rows, int* ncols);
int svm_table_column_type(const char* id, int col, int32_t* type, int64_t* val_size);
/* Zero-copy view `id` of rows [row0, row0+rows) of table `base_id` (a fact slice). */
int svm_table_view(const char* id, const char* base_id, uint64_t row0, uint64_t rows);
int svm_table_drop(const char* id);
int svm_table_exists(const char* id);

/* ---- synthetic star-schema generator (tables generated directly in HBM) ---- */
typedef struct svm_gen_params {
  uint64_t n_cust, n_supp, n_part; /* foreign-key ranges for fact tables */
  double zipf_s;                   /* FACT3 Zipf exponent (1.1) */
} svm_gen_params;
/* kind: svm_gen_kind in shadowvm_amd/csrc/datagen.h (0 DATE ... 8 FACT3). */
int svm_gen_table(const char* id, int kind, uint64_t rows, uint64_t seed, const svm_gen_params* p);
/* The same rows [row0, row0+n) on the host into caller column buffers. */
int svm_gen_host(int kind, uint64_t seed, const svm_gen_params* p, uint64_t row0, uint64_t n,
                 void* const* cols, int threads);
/* Schema of a generated table kind: fills up to 8 columns, returns the count. */
int svm_gen_schema(int kind, int32_t* types, int64_t* val_sizes);

/* ---- pipelines (ref RunPipelineCMD, service/service.cu:25-63) ---- */
typedef struct svm_result svm_result;
/* Decode serialized proto3 PipelineInfo bytes, execute on the GPU, cache the
 * output table under PipelineInfo.id and (if out != NULL) return it on the host. */
int svm_pipeline_run(const uint8_t* pipelineinfo_pb, size_t len, svm_result** out);

/* svm_pipeline_run without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Runs a ReduceBy (fact) pipeline
 * and writes its result into caller device buffers: out_cols_device[0..K) are K Int32
 * key columns, [K..K+M) M Int64 sum columns (ncols = K + M; a scalar has K = 0 and
 * one row), each max_rows rows; *rows receives the number of result rows (also when
 * it exceeds max_rows, which fails with SVM_ERR_INVALID).  Calls no hipMalloc /
 * hipFree / allocating routine: the run's temporaries (join maps, group tables,
 * emission offsets) come from a static 64 MiB device scratch array per GPU, and a
 * run that needs more fails with SVM_ERR_NO_MEMORY.  Everything a first run builds
 * and caches (key ranges, dictionaries, join hash tables, compiled kernels, pinned
 * host words) must exist already: run the same pipeline shape once with
 * svm_pipeline_run; otherwise the call fails with SVM_ERR_STATE and allocates
 * nothing.  The output object is not cached under PipelineInfo.id, no autotuning
 * trial runs and the run is untimed.  Calls are serialised with each other and with
 * svm_pipeline_run (one engine stream); the scratch array keeps the last call's
 * contents between calls. */
int svm_pipeline_run_noalloc(const uint8_t* pipelineinfo_pb, size_t len, void* const* out_cols_device, int ncols,
                             uint64_t max_rows, uint64_t* rows);

/* Result access.  Group-by results: K Int32 key columns then M Int64 sums,
 * ordered by key; scalar: one row of M Int64; select: the sink schema's
 * projected columns. */
int svm_result_kind(const svm_result* r); /* 0 select, 1 group-by, 2 scalar */
int64_t svm_result_rows(const svm_result* r);
int svm_result_cols(const svm_result* r);
int svm_result_col_type(const svm_result* r, int col, int32_t* type, int64_t* val_size);
const void* svm_result_col(const svm_result* r, int col);
/* Timing of the last run of this result: build, scan and finish, in ms (GPU events). */
int svm_result_timing(const svm_result* r, double* build_ms, double* scan_ms, double* finish_ms);
/* Everything the accessors above return, in one call (no reference counterpart: it
 * saves a binding one foreign call per field when many results are read). */
typedef struct svm_result_desc {
  int32_t kind;               /* as svm_result_kind */
  int32_t ncols;              /* <= 24 (MaxColumnNum, com/cpp/define.hpp:25) */
  int64_t rows;
  int32_t type[24];
  int64_t val_size[24];
  const void* col[24];        /* host column data, rows * val_size bytes each */
  double build_ms, scan_ms, finish_ms;
} svm_result_desc;
int svm_result_describe(const svm_result* r, svm_result_desc* d);
void svm_result_free(svm_result* r);

/* The same for n results at once (no reference counterpart; for bindings that read
 * a whole batch): describe fills d[i] for each non-NULL rs[i] (kind -1 for NULL);
 * copy writes every column of every non-NULL result into dst, results in order,
 * columns in order, rows * val_size bytes each, each column starting at the next
 * multiple of 8 bytes; free releases them all. */
int svm_results_describe(int n, svm_result* const* rs, svm_result_desc* d);
int svm_results_copy(int n, svm_result* const* rs, void* dst, uint64_t dst_bytes);
void svm_results_free(int n, svm_result* const* rs);

/* ---- batched execution: many star-join queries fused into one shared scan ----
 * plans: dimension and fact pipelines of all queries, in submission order.  Fact
 * pipelines whose operators fit the shared scan (range and StrCmp equality atoms,
 * bare-column projections, equi-joins on unique int32 keys, 0-3 group keys, one
 * metric a, a-b, a+b or a*b over int32 fact columns) run fused; their
 * dimension pipelines are absorbed (outs[i] is NULL for an absorbed one and
 * its object is not cached); every other plan runs as svm_pipeline_run would.
 * Returns the number of fused fact pipelines (>= 0) or a negative status. */
int svm_batch_run(int n, const uint8_t* const* plans, const size_t* lens, svm_result** outs);

/* ---- serving: continuous batching against device-resident tables ----
 * A query is submitted as its plans (dimension pipelines, then the fact pipeline
 * last).  Queries waiting for admission are taken as HBM budget (maps, group
 * tables, descriptors) and worker slots free up, fused into one shared scan per
 * batch, and run on `workers` streams so batches overlap on the device.
 * mode 0: continuous batching; mode 1: one query per launch through the plain
 * kernels (the comparison baseline).  max_batch caps the queries of one batch. */
int svm_server_start(int workers, uint64_t budget_bytes, int max_batch, int mode);
int svm_server_submit(const uint8_t* const* plans, const size_t* lens, int n, uint64_t* ticket);
/* Waits up to timeout_ms (<0: forever) for the fact pipeline's result; returns
 * SVM_ERR_BUSY on timeout.  out may be NULL.  The ticket is released on success. */
int svm_server_wait(uint64_t ticket, int timeout_ms, svm_result** out);
/* svm_server_wait that also reports the query's server-side latency: ms from its
 * submission to its admission into a batch, and to its result being ready.
 * Either pointer may be NULL. */
int svm_server_wait_timed(uint64_t ticket, int timeout_ms, svm_result** out, double* queued_ms,
                          double* total_ms);
/* Queue length, admitted queries, completed queries and HBM bytes in use. */
int svm_server_stats(uint64_t* queued, uint64_t* running, uint64_t* done, uint64_t* bytes_in_use);
int svm_server_stop(void);

#ifdef __cplusplus
}
#endif

#endif /* SVM_H_ */
