// codegen.h — HIP source generation for one pipeline shape (compiled by hipRTC).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace svm {

// What a join's packed dense map stores per key slot (0 always means "no row").
enum PayloadKind {
  PAY_EXIST = 0,  // 1: a dimension row with this key passed the dimension pipeline
  PAY_CODE = 1,   // code+1 of the single group-key component this dimension supplies
  PAY_ROW = 2     // row+1 of the first dimension row with this key
};

// How a join finds its dimension row.
enum JoinKind {
  JOIN_MAP = 0,    // Int32 keys: dense map over the key range
  JOIN_PROBE = 1,  // the reference's hash table, Probe<T> over the key's bytes
  JOIN_PROBE_STR = 2  // the reference's hash table, Probe<char> (StrHash, StrCmp)
};

struct JoinShape {
  int robj = 1;      // ref object (D4: join_index is addressed by this object)
  int lcol = 0;      // source (fact) column holding the foreign key
  int payload = PAY_EXIST;
  int wbits = 1;     // 1,2,4,8,16 or 32 bits per key slot
  bool lds = false;  // map staged in LDS at kernel start
  // The scan tests a 1-bit existence map first (stored after the packed words, at
  // word a.iv[4j+2]) and reads the packed payload only for rows that pass every
  // join: a large code map then stays out of the probe chain.
  bool defer = false;
  // a deferred map's existence bits staged in LDS at word a.iv[4j+3]
  bool bits_lds = false;
  int kind = JOIN_MAP;
  int rcol = 0;      // probes: the dimension's key column (read as DCOL<robj>_<rcol>)
  int ktype = 2;     // probes: key type (ValType) and width in bytes
  int ksize = 4;
};

enum KeySource { KEY_DIMCODE = 0, KEY_DIMROW = 1, KEY_FACT = 2 };
struct KeyShape {
  int src = KEY_DIMCODE;
  int join = 0;  // KEY_DIMCODE / KEY_DIMROW: join position supplying the row
  int obj = 0;
  int col = 0;   // KEY_FACT: source column
};

// AGG_SORT: group spaces a dense table cannot hold.  The key codes are packed into
// one 64-bit key (key q at bit a.iv[32+4q+3]), pre-aggregated in an LDS hash table
// and appended as (key, sums) entries to HBM (a.out[3] keys, a.out[4] sums, a.out[5]
// entry counter, capacity a.iv[46] | a.iv[47] << 32) for a sort + reduce by key.
enum AggMode { AGG_SCALAR = 0, AGG_DENSE = 1, AGG_HASH = 2, AGG_SORT = 3 };

struct KernelShape {
  int block = 256;
  std::vector<int> fact_i32;          // Int32 source columns held in registers
  std::string filter = "true";        // lowered conjunction of Filter predicates
  std::vector<JoinShape> joins;
  std::map<std::pair<int, int>, int> dcol_slot;  // (obj,col) -> ptr slot (ROW payload)
  std::vector<int> fact_other;        // non-Int32 source columns read by expressions
  // reduce
  std::vector<KeyShape> keys;
  std::vector<std::string> metrics;      // lowered metric expressions
  std::vector<std::string> metric_ctype; // C type of the declared metric column
  int agg = AGG_SCALAR;
  int hash_bits = 11;                    // LDS hash table of 2^hash_bits slots
  int hash_probes = 16;
  int unroll = 1;                        // int4 groups per thread per iteration
  bool prefetch = false;                 // load the next iteration's groups before processing
  bool flat = false;                     // branch-free filter + joins for the four rows, then aggregate
  // Staged column loads (flat form; empty: every column is loaded up front).  The key
  // column of join j is loaded in tier join_tier[j] (-1: with the last columns); a
  // lane loads the columns of tier t > 0 only when one of its four rows passed every
  // join of the tiers before, and the columns of no tier (metrics, group keys, filter
  // columns, the keys of untiered joins) after the last tier.
  std::vector<int> join_tier;
  // staged, compacting form: tier 0 is streamed by every lane; each wave then hands
  // the rows that pass it out one per lane, and those rows' later tiers, filter and
  // aggregation run row by row from 4-byte gathers
  bool compact = false;
  // compacting form: the one tier-0 column, when it is packed[0], is streamed from
  // its narrow copy (a.ptr[SVM_SLOT_PACK], minimum a.iv[48]; blocks of 512 rows as a
  // 16-bit plane and a plane of pack_high = 0, 2 or 8 high bits a row): a lane reads 8
  // rows per 128-bit load plus one load of 8 x pack_high bits, `unroll` groups of 8
  // rows per iteration.  Later tiers gather the Int32 columns.
  std::vector<int> packed;
  int pack_high = 8;
  // select
  std::vector<std::string> proj;                      // lowered projection expressions
  std::vector<std::pair<int, long long>> proj_type;   // (ValType, val_size) of outputs
};

// Generated source text and its source_hash.
struct GeneratedSource {
  std::string text;
  uint64_t hash = 0;
};
// gen_scan_source / gen_select_source memoised per kernel shape (the text of a shape
// never changes; constants are launch parameters).
std::shared_ptr<const GeneratedSource> scan_source(const KernelShape& s, const char* entry = "svm_scan");
std::shared_ptr<const GeneratedSource> select_source(const KernelShape& s);

// entry: the kernel's symbol (the autotuner's trial launches use "svm_scan_tune",
// so that profiles tell them from the query's own scans)
std::string gen_scan_source(const KernelShape& s, const char* entry = "svm_scan");
std::string gen_select_source(const KernelShape& s);

// The fact column the compacting form of `s` can stream from a narrow copy: its one
// tier-0 column, when it has one and the filter is applied later; else empty.
std::vector<int> compact_packable(const KernelShape& s);
// C++ source of the host engine's task function for the same pipeline shape:
//   extern "C" void svm_host_task(const SvmArgs*, long long r_lo, long long r_hi, SvmHostOut*)
// (a select when s.proj is set, else the reduce of s.agg).
std::string gen_host_source(const KernelShape& s);
std::string device_runtime_text();
std::string host_runtime_text();

}  // namespace svm
