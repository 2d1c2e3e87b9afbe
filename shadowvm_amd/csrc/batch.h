// batch.h — descriptors of the batched (shared-scan) execution path.
//
// A batch holds many star-join queries over one resident fact table.  Each query
// is lowered from its reference plans (dimension filter+map pipelines and the fact
// filter/join/reduce pipeline) into a fixed-size descriptor that one precompiled,
// parameter-driven kernel interprets, so a batch of any shape runs without
// compiling anything and its launch sequence can be captured once in a hipGraph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace svm {
namespace bk {

constexpr int kMaxPred = 4;     // range predicates on fact columns per query
constexpr int kMaxJoin = 4;     // joins per query
constexpr int kMaxKey = 3;      // group-key components per query
constexpr int kMaxDimPred = 3;  // atoms of one dimension filter
constexpr int kMaxCols = 8;     // fact columns the shared scan can read
constexpr int kDenseG = 2048;   // groups held in a dense LDS table
constexpr int kHashSlots = 1024;

enum DimPredKind { DP_RANGE = 0, DP_STREQ = 1, DP_STRNE = 2 };

struct DimPred {
  const char* col;  // row 0 of the dimension column
  int kind;
  int type;         // ValType of the column (range atoms)
  int size;         // val_size
  int lo, hi;       // DP_RANGE: lo <= value <= hi
  int n;            // StrCmp length
  char lit[48];     // NUL-padded literal
};

// Builds the packed existence / code map of one (query, join).
struct MapBuild {
  const int32_t* keys;  // dimension key column (row 0)
  unsigned long long n;
  int kmin;
  unsigned range;
  int npred;
  DimPred pred[kMaxDimPred];
  const int32_t* codes;  // global dictionary code per dimension row, or null
  int wbits;
  unsigned* map;         // zeroed packed words
};

enum MetricOp { MOP_COL = 0, MOP_SUB = 1, MOP_ADD = 2, MOP_MUL = 3 };
enum AggKind { BAGG_SCALAR = 0, BAGG_DENSE = 1, BAGG_HASH = 2 };

struct Query {
  long long row0, row1;  // slice of the base fact table
  int npred;
  int pcol[kMaxPred];
  int plo[kMaxPred], phi[kMaxPred];
  int njoin;
  int jcol[kMaxJoin];
  const unsigned* jmap[kMaxJoin];
  int jkmin[kMaxJoin];
  unsigned jrange[kMaxJoin];
  int jw[kMaxJoin];
  int jcode[kMaxJoin];        // 1: payload is code+1 of a key component
  unsigned jstride[kMaxJoin]; // group-id stride of that component
  int mop, ma, mb;
  int agg;
  unsigned G;
  unsigned long long* sums;   // G int64 sums
  unsigned* flags;            // bit per group
};

struct Item {
  int tile;
  int q;
};

// Per-query group emission.
struct Emit {
  const unsigned* flags;
  const unsigned long long* sums;
  unsigned G;
  int K;
  unsigned n[kMaxKey], stride[kMaxKey];
  const int32_t* dict[kMaxKey];
  unsigned long long block0;  // first emission block of this query
  unsigned long long nblocks;
  unsigned long long out0;    // first output row of this query in the arena (filled after counting)
};

void launch_map_build(const MapBuild* d_builds, int nbuild, unsigned long long max_rows, hipStream_t s);
void launch_shared_scan(const Query* d_q, const Item* d_items, unsigned nitems, unsigned* d_counter,
                        const void* const* cols, long long tile_rows, int grid, hipStream_t s);
void launch_emit_count(const Emit* d_e, int nq, unsigned long long total_blocks, unsigned long long* d_counts,
                       unsigned long long* d_offsets, unsigned long long* d_totals, hipStream_t s);
void launch_emit_rows(const Emit* d_e, int nq, unsigned long long total_blocks,
                      const unsigned long long* d_offsets, int32_t* keys_arena, int64_t* sums_arena,
                      unsigned long long arena_rows, hipStream_t s);

}  // namespace bk
}  // namespace svm
