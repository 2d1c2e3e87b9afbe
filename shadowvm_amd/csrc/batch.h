// batch.h — descriptors of the batched (shared-scan) execution path.
//
// A batch holds many star-join queries over one resident fact table.  Each query
// is lowered from its reference plans (dimension filter+map pipelines and the fact
// filter/join/reduce pipeline) into fixed-size descriptors that precompiled,
// parameter-driven kernels interpret, so a batch of any shape runs without
// compiling anything.
//
// Joins of the batch are grouped into slots: one slot per (dimension table, key
// column, fact column).  Every row of a slot's dimension carries a profile — the
// packed dictionary codes of the dimension columns that the batch's queries filter
// or group on — deduplicated into a compact id per key.  A fact row then needs one
// lookup per slot (its key's profile id) whatever the number of queries: the
// profile's 64-bit mask says which queries' dimension filters the row passes and
// its packed codes give every query's group-key components.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "refhash.h"

namespace svm {
namespace bk {

constexpr int kMaxPred = 4;          // range predicates on fact columns per query
constexpr int kMaxJoin = 4;          // joins per query
constexpr int kMaxKey = 3;           // group-key components per query
constexpr int kMaxDimPred = 3;       // atoms of one dimension filter
constexpr int kMaxCols = 6;          // fact columns the shared scan can read
constexpr int kMaxBatch = 64;        // queries of one shared scan (bits of a profile mask)
constexpr int kMaxSlot = 6;          // slots of one shared scan
constexpr int kMaxAttr = 6;          // dimension columns one slot's profiles encode
constexpr int kMaxProfiles = 4096;   // profiles of all slots of one scan (staged in LDS)
constexpr int kMaxMetric = 4;        // distinct metric expressions of one scan
constexpr int kMaxFactPred = 128;    // (query, column, range) entries of one scan
constexpr int kMaxSegBounds = 2 * kMaxFactPred;  // interval bounds of all fact predicates
constexpr int kMaxTerms = kMaxBatch * kMaxJoin;
constexpr unsigned kMaxProfilesPerSlot = 0xFFFEu;  // profile ids and the no-row id P fit 16 bits

enum DimPredKind { DP_RANGE = 0, DP_STREQ = 1, DP_STRNE = 2 };
enum MetricOp { MOP_COL = 0, MOP_SUB = 1, MOP_ADD = 2, MOP_MUL = 3 };

// One dimension atom evaluated on a profile: the attribute's code is
// (raw >> shift) & ((1 << bits) - 1); a range atom compares the code's value,
// a string atom the bytes of the code's first dimension row.
struct MaskAtom {
  int kind;
  int shift, bits;
  int lo, hi;               // DP_RANGE: lo <= value <= hi
  int n;                    // StrCmp length
  int size;                 // string column width
  const int32_t* dict;      // DP_RANGE: value of each code
  const char* col;          // string atoms: the column (row 0)
  const int32_t* rep;       // string atoms: first row of each code
  char lit[48];             // NUL-padded literal
};

// One query's join on a slot: the query passes a profile when all atoms pass.
struct MaskTerm {
  int q;
  int natom;
  MaskAtom atom[kMaxDimPred];
};

struct MaskSlot {
  unsigned P;                   // profiles
  const int32_t* praw;          // packed codes of each profile
  unsigned long long* mask;     // out: queries each profile passes
  int t0, t1;                   // its terms
};

struct MaskArgs {
  MaskSlot slots[kMaxSlot];
  const MaskTerm* terms;
  int nslot;
};

// A slot as the shared scan reads it.
struct Slot {
  int fcol;                     // fact column holding the join key
  int kmin;
  unsigned range;
  unsigned P;
  const void* pkey;             // profile of each key offset (P: no dimension row), uint8
  int narrow;                   //   when narrow (P < 255) else uint16; entry `range` holds P
  const unsigned long long* mask;
  const int32_t* praw;
  unsigned long long joined;    // queries that join this slot
  unsigned* bits;               // per key offset: some query of the batch passes it (built per batch)
  int lds;                      // first profile of this slot in the LDS staging arrays (P + 1 entries)
  // Variants with LDS maps (assigned per pass by assign_lds_maps): the key-to-profile
  // map packed at 1 << lmap_log2 bits per entry starts at word lmap of the scan's LDS
  // map area, or lmap is -1; the key bits start at word lbits there, or lbits is -1.
  int lmap, lmap_log2, lbits;
};

// Per-query descriptor read by every lane that aggregates a row for the query.
struct QDesc {
  unsigned key[kMaxKey];        // lds raw offset | slot << 13 | shift << 16 | bits << 21
  unsigned stride[kMaxKey];
  int K;
  int metric;                   // index into the scan's distinct metrics
  unsigned sum_off, flag_off;   // the query's group table in the batch arenas
  unsigned G;                   // groups of the query
  unsigned dense;               // dense-accumulator variants: first LDS slot of the query's
                                //   G groups, or kNoDense (its pairs go to the HBM tables)
};
constexpr unsigned kNoDense = 0xFFFFFFFFu;

struct FactPred {
  int q;
  int lo, hi;
};

// A slot whose dimension table drops inserts in the reference's hash table
// (core/pipeline.cu:88-133): the clusters of its inserts that drop something.  A
// query's dimension object is an in-order subset of the table's rows, so it can drop
// a row only inside one of these clusters, and only where it keeps at least 11 of
// the cluster's rows (a drop needs 10 slots taken by earlier rows of the cluster).
struct DropCheck {
  const unsigned* start;           // cluster c holds entries [start[c], start[c+1])
  const int* key;                  // key of each entry (row order inside a cluster)
  const unsigned* home;            // home slot of each entry
  unsigned clusters;
  int kmin;
  unsigned range, P;
  const void* pkey;                // profile per key offset (uint8 when narrow, else uint16)
  int narrow;
  const unsigned long long* mask;  // query mask of each profile < P
  unsigned long long joined;       // queries joining through this slot
};
constexpr int kCheckSpan = k::kReplaySpan;  // widest cluster (in slots) a check replays

struct ScanArgs {
  const void* cols[kMaxCols];
  unsigned qcols[kMaxBatch];    // fact columns each query reads
  int nq, nslot, nmetric;
  int pred0[kMaxCols + 1];      // FactPreds of column c are pred[pred0[c] .. pred0[c+1])
  FactPred pred[kMaxFactPred];
  int mop[kMaxMetric], ma[kMaxMetric], mb[kMaxMetric];
  Slot slot[kMaxSlot];
  unsigned nprof;               // profiles staged in LDS (sum over slots)
  int use_bits;                 // test each slot's key bit before reading the profile
  const QDesc* qd;
  const long long* qrow;        // row0, row1 of each query's slice of the base table
  unsigned long long* sums;
  unsigned* flags;
  unsigned long long sum_words, flag_words;
  unsigned long long total_blocks;  // emission blocks of all queries
  long long row_lo, tile_rows;  // tile t covers [row_lo + t*tile_rows, ...) clipped to row_hi
  long long row_hi;
  unsigned ntiles;
  unsigned* counter;
  // The fact predicates as value segments per column: the predicates' interval
  // bounds of column c, ascending and distinct, are segb[seg0[c] .. seg0[c+1]); a value
  // lies in segment j = the number of those bounds <= it, whose query mask (the
  // queries all of whose predicates on c hold there) is segm[seg0[c] + c + j].  colq[c]:
  // queries with a predicate on c; segsteps[c]: steps of the search over its bounds.
  int seg0[kMaxCols + 1];
  int segb[kMaxSegBounds];
  unsigned long long segm[kMaxSegBounds + kMaxCols];
  unsigned long long colq[kMaxCols];
  int segsteps[kMaxCols];
  // Staged variants: the columns in tier0 are loaded for every row group, slots
  // 0..nslot0-1 (whose key columns they hold) tested first, and the other columns
  // loaded only for lanes with a row that still passes some query.
  unsigned tier0;
  int nslot0;
  const DropCheck* chk;         // slots to check for dropped inserts
  int nchk;
  unsigned long long* dropq;    // bit q: query q's dimension object drops an insert
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

// Profile construction (cached per dimension table, key column and attribute set).
struct RawArgs {
  const int32_t* codes[kMaxAttr];
  int shift[kMaxAttr];
  int nattr;
};
void launch_profile_raw(const RawArgs& a, uint64_t n, int32_t* raw, hipStream_t s);
// pkey[keys[r] - kmin] = pcode[r] as uint8 (narrow) or uint16 (pkey initialised to P).
void launch_profile_scatter(const int32_t* keys, uint64_t n, int32_t kmin, const int32_t* pcode, void* pkey,
                            int narrow, hipStream_t s);
// *flag = 1 if a row's bytes differ from those of its code's first row.
void launch_string_exact(const char* col, int size, uint64_t n, const int32_t* codes, const int32_t* rep,
                         unsigned* flag, hipStream_t s);

// The launch sequence of one pass.  Grids and arguments do not depend on the
// batch, so the sequence can be captured once and replayed.
void launch_pass_clear(const ScanArgs* d_args, hipStream_t s);
void launch_profile_masks(const MaskArgs* d_margs, hipStream_t s);
void launch_slot_bits(const ScanArgs* d_args, hipStream_t s);
// Sets dropq bit q for every query whose dimension object on a checked slot drops an
// insert (replays the insert order of the clusters where the query keeps >= 11 rows).
void launch_drop_check(const ScanArgs* d_args, hipStream_t s);
// Shared-scan variants (workgroup size x LDS accumulator size) the autotuner picks from.
int scan_variants();
constexpr int kDefaultVariantMask32 = 11;  // kVariants: 1,024 threads, dense, LDS maps, queue, 32-bit masks
int scan_variant_block(int v);
int scan_variant_hash_slots(int v);
bool scan_variant_parallel(int v);  // slot lookups issued together rather than gated in turn
bool scan_variant_mask32(int v);    // 32-bit query masks (scans of at most 32 queries)
bool scan_variant_prefetch(int v);  // next row group's loads in flight during this one's lookups
bool scan_variant_staged(int v);    // ScanArgs::tier0 columns first, the rest for passing lanes only
// Dense-accumulator variants: a query with few groups owns scan_variant_hash_slots(v)
// LDS slots' worth of a per-workgroup dense table (QDesc::dense, assigned per pass by
// assign_dense); the others add straight into their HBM tables.
bool scan_variant_dense(int v);
void assign_dense(QDesc* qd, int nq, unsigned slots);
// Variants that stage small key-to-profile maps and key bits in LDS: the maps go to
// the slots with the smallest packed maps while they fit, the bits to the rest.
bool scan_variant_lds_maps(int v);
void assign_lds_maps(ScanArgs* A, int variant, unsigned nprof_cap);
size_t shared_scan_lds(int variant, unsigned nprof_cap);
void prepare_shared_scan();  // LDS limits of every variant; call before any capture
void launch_shared_scan(const ScanArgs* d_args, int variant, unsigned nprof_cap, int grid, hipStream_t s);
void launch_emit_count(const Emit* d_e, const ScanArgs* d_args, unsigned long long* d_counts,
                       unsigned long long* d_offsets, unsigned long long* d_totals, hipStream_t s);
void launch_emit_rows(const Emit* d_e, int nq, unsigned long long total_blocks,
                      const unsigned long long* d_offsets, int32_t* keys_arena, int64_t* sums_arena,
                      unsigned long long arena_rows, hipStream_t s);

}  // namespace bk
}  // namespace svm
