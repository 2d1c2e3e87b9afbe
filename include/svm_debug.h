/*
 * svm_debug.h — inspection entry points of libsvm.so (the first two need no GPU).
 *
 * These have no counterpart in the reference; they expose what the reference
 * only prints (the decoded plan, core/pipeline.cu:11-80, and the generated code,
 * test/test.hpp:106-114), so the boundary and the code generator can be checked
 * on a machine without a GPU, and what it does silently (the inserts its join hash
 * tables drop).
 */
#ifndef SVM_DEBUG_H_
#define SVM_DEBUG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Decode PipelineInfo bytes and write the operator chain as JSON into buf:
 * id, pre ids, filter predicates with their lowered HIP text, select / metric
 * lists, reduce and join keys, literal parameters, output schema. */
int svm_debug_describe_plan(const uint8_t* pipelineinfo_pb, size_t len, char* buf, size_t cap);

/* Generate representative kernels (scalar / dense / hash / row-id scans and a
 * select) and compile them with hipRTC for gfx950; writes a report, returns 0
 * when all compile. */
int svm_debug_jit_selftest(char* buf, size_t cap);

/* Copy column `src_col` of table `src_id` into column `dst_col` of `dst_id` (same
 * rows and width) with a plain 128-bit streaming copy kernel (svm_calib_copy).
 * A kernel of known traffic (read n, write n bytes) to calibrate HBM counters. */
int svm_debug_copy_column(const char* dst_id, int dst_col, const char* src_id, int src_col);

/* What HBM gives a plain read of a scan's columns: Int32 columns cols[0..ncols) (<= 8)
 * of table `id` streamed with non-temporal 128-bit loads and nothing else
 * (svm_read_probe), blocks_per_cu workgroups of `block` threads per CU; *best_ms
 * receives the fastest of `reps` launches (bytes = rows / 4 * 16 * ncols). */
int svm_debug_read_probe(const char* id, int ncols, const int* cols, int block, int blocks_per_cu, int reps,
                         double* best_ms);

/* Inserts the reference's join hash table would drop (BuildHashTable,
 * com/cpp/com.hpp:168-174: Hash of the key bytes % slots, at most 10 linear probes,
 * silent drop) over every row of Int32 column `col` of table `id`, inserted in row
 * order into `slots` slots (0: HashRate x the table's bulk capacity).  Computed on
 * the GPU, as the joins do. */
int svm_debug_hash_drops(const char* id, int col, uint64_t slots, uint64_t* dropped);

/* The autotuners' record as JSON {"scan": [...], "batch": [...]}: for every tuned
 * shape the chosen launch configuration and each trial's time in ms.  Trials are
 * cross-checked on the device against the first trial's group tables; a negative time
 * marks a trial that differed (never chosen; "rejected" counts them).  reset != 0
 * clears both caches afterwards.  Environment controls read by the tuners:
 * SVM_SCAN_VARIANT=<i> / SVM_BATCH_VARIANT=<v>,<tile>,<bits> force one candidate,
 * SVM_TUNE_MIN_ROWS / SVM_BATCH_TUNE_MIN_ROWS lower the size from which they tune,
 * SVM_AUTOTUNE=0 turns tuning off. */
int svm_debug_tune_report(char* buf, size_t cap, int reset);

/* How many launch configurations each autotuner chooses from: the fact scan's
 * candidates (SVM_SCAN_VARIANT=0..n-1) and the shared scan's kernel variants (the
 * first field of SVM_BATCH_VARIANT).  Needs no GPU. */
int svm_debug_variant_counts(int* scan_candidates, int* batch_variants);

/* The launch configuration of the last fact scan of pipeline `pipeline_id`, as text:
 * "<threads>x<groups per thread> lds-maps|l2-maps [prefetch] [flat]" and, for a scan
 * with staged column loads, " staged(...)", " chained-tiers(...)" or
 * " compacted-tiers(...)" listing each join's load tier in plan order (-1: loaded
 * with the metric and key columns after the last tier).  SVM_ERR_NOT_FOUND when no
 * scan of that pipeline has run. */
int svm_debug_scan_config(const char* pipeline_id, char* buf, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* SVM_DEBUG_H_ */
