// selftest.cpp — compile representative generated kernels with hipRTC without a
// GPU (code generation + compiler check; also warms the code-object cache).
#include <cstring>

#include "codegen.h"
#include "host_engine.h"
#include "jit.h"
#include "svm_internal.h"

namespace svm {

namespace {
// The freshly generated source, checked against the memoised one for the same shape
// (the shapes below differ from each other in one field at a time, so a memo key that
// missed a field would return another shape's text here).
std::string scan_checked(const KernelShape& s) {
  std::string t = gen_scan_source(s);
  if (scan_source(s)->text != t || scan_source(s)->hash != source_hash(t))
    fail(SVM_ERR_JIT, "memoised scan source differs from the generated one");
  return t;
}
std::string select_checked(const KernelShape& s) {
  std::string t = gen_select_source(s);
  if (select_source(s)->text != t || select_source(s)->hash != source_hash(t))
    fail(SVM_ERR_JIT, "memoised select source differs from the generated one");
  return t;
}
}  // namespace

std::vector<std::pair<std::string, std::string>> selftest_sources() {
  std::vector<std::pair<std::string, std::string>> out;
  {  // single join, scalar sum, map in LDS
    KernelShape s;
    s.fact_i32 = {0, 1, 2, 3};
    s.filter = "(f2)>=PI(0)&&(f2)<=PI(1)&&(f1)<PI(2)";
    JoinShape j; j.robj = 1; j.lcol = 0; j.payload = PAY_EXIST; j.wbits = 1; j.lds = true;
    s.joins = {j};
    s.metrics = {"(f3)*(f2)"};
    s.metric_ctype = {"int"};
    s.agg = AGG_SCALAR;
    out.push_back({"scan_scalar", scan_checked(s)});
  }
  {  // three joins, dense LDS group table
    KernelShape s;
    s.fact_i32 = {0, 1, 2, 3};
    JoinShape a; a.robj = 1; a.lcol = 0; a.payload = PAY_CODE; a.wbits = 4; a.lds = true;
    JoinShape b; b.robj = 2; b.lcol = 1; b.payload = PAY_CODE; b.wbits = 16;
    JoinShape c; c.robj = 3; c.lcol = 2; c.payload = PAY_EXIST; c.wbits = 1;
    s.joins = {a, b, c};
    KeyShape k0; k0.src = KEY_DIMCODE; k0.join = 0; k0.obj = 1;
    KeyShape k1; k1.src = KEY_DIMCODE; k1.join = 1; k1.obj = 2;
    s.keys = {k0, k1};
    s.metrics = {"(f3)"};
    s.metric_ctype = {"int"};
    s.agg = AGG_DENSE;
    out.push_back({"scan_dense", scan_checked(s)});
  }
  {  // four joins, LDS hash group table, two int4 groups per iteration
    KernelShape s;
    s.fact_i32 = {0, 1, 2, 3, 4, 5};
    s.unroll = 2;
    JoinShape a; a.robj = 1; a.lcol = 0; a.payload = PAY_CODE; a.wbits = 2; a.lds = true;
    JoinShape b; b.robj = 2; b.lcol = 1; b.payload = PAY_CODE; b.wbits = 8;
    JoinShape c; c.robj = 3; c.lcol = 2; c.payload = PAY_CODE; c.wbits = 8;
    JoinShape d; d.robj = 4; d.lcol = 3; d.payload = PAY_EXIST; d.wbits = 1;
    s.joins = {a, b, c, d};
    KeyShape k0; k0.src = KEY_DIMCODE; k0.join = 0; k0.obj = 1;
    KeyShape k1; k1.src = KEY_DIMCODE; k1.join = 1; k1.obj = 2;
    KeyShape k2; k2.src = KEY_DIMCODE; k2.join = 2; k2.obj = 3;
    s.keys = {k0, k1, k2};
    s.metrics = {"(f4)-(f5)"};
    s.metric_ctype = {"int"};
    s.agg = AGG_HASH;
    s.hash_bits = 10;
    out.push_back({"scan_hash", scan_checked(s)});
    s.unroll = 1;
    s.prefetch = true;
    s.block = 512;
    out.push_back({"scan_hash_prefetch", scan_checked(s)});
    s.prefetch = false;
    s.unroll = 2;
    s.flat = true;
    out.push_back({"scan_hash_flat", scan_checked(s)});
    // the large code maps probed through their existence bits
    s.joins[1].defer = true;
    s.joins[2].defer = true;
    out.push_back({"scan_hash_flat_deferred", scan_checked(s)});
    // staged loads: the key columns of joins 1 and 2 first, the rest for lanes with a
    // passing row (one tier; then a tier each, with the first tier prefetched)
    s.join_tier = {-1, 0, 0, -1};
    out.push_back({"scan_staged", scan_checked(s)});
    s.join_tier = {2, 0, 1, -1};
    s.joins[1].bits_lds = true;  // existence bits of a deferred map in LDS
    s.unroll = 1;
    s.prefetch = true;
    s.filter = "(f4)>PI(0)&&(f1)!=PI(1)";
    out.push_back({"scan_staged_chain", scan_checked(s)});
    s.compact = true;  // the rows passing tier 0 compacted per wave
    s.prefetch = false;
    s.unroll = 4;
    out.push_back({"scan_compacted", scan_checked(s)});
    // its tier-0 key column streamed from a narrow copy (24, 18 and 16 bits a row),
    // four groups of 8 rows a lane
    s.packed = compact_packable(s);
    if (s.packed.size() != 1 || s.packed[0] != 1) fail(SVM_ERR_JIT, "compact_packable: unexpected columns");
    s.unroll = 4;
    out.push_back({"scan_compacted_packed", scan_checked(s)});
    s.pack_high = 2;
    out.push_back({"scan_compacted_packed18", scan_checked(s)});
    s.pack_high = 0;
    out.push_back({"scan_compacted_packed16", scan_checked(s)});
    s.pack_high = 8;
    s.packed.clear();
    s.compact = false;
    s.join_tier.clear();
    s.joins[1].bits_lds = false;
    s.prefetch = false;
    s.filter = "true";
    s.flat = false;
    out.push_back({"scan_hash_deferred", scan_checked(s)});
  }
  {  // row-id payload, dimension attribute in the metric, source-column key
    KernelShape s;
    s.fact_i32 = {0, 1, 3};
    JoinShape a; a.robj = 1; a.lcol = 0; a.payload = PAY_ROW; a.wbits = 32;
    s.joins = {a};
    s.dcol_slot[{1, 1}] = 32;
    KeyShape k0; k0.src = KEY_DIMROW; k0.join = 0; k0.obj = 1; k0.col = 2;
    KeyShape k1; k1.src = KEY_FACT; k1.col = 3;
    s.keys = {k0, k1};
    s.metrics = {"(*(const int*)(DCOL1_1+(long long)JI1*4))*(f1)"};
    s.metric_ctype = {"long long"};
    s.agg = AGG_HASH;
    out.push_back({"scan_row", scan_checked(s)});
    s.flat = true;
    s.filter = "StrCmp(((const char*)(FCOL2+(long long)i*10)),PS(0),PI(0))!=PI(1)";
    s.fact_other = {2};
    out.push_back({"scan_row_flat", scan_checked(s)});
    // a group space too wide for a table: packed 64-bit keys, sort + reduce; two metrics
    s.agg = AGG_SORT;
    s.metrics.push_back("(f0)-(f1)");
    s.metric_ctype.push_back("int");
    out.push_back({"scan_sort_flat", scan_checked(s)});
    s.flat = false;
    out.push_back({"scan_sort", scan_checked(s)});
  }
  {  // probe joins on the reference's hash table: an Int64 key and an Int32 key, grouped
    KernelShape s;
    s.fact_i32 = {1, 3};
    s.fact_other = {0};
    JoinShape a; a.robj = 1; a.lcol = 0; a.payload = PAY_ROW; a.wbits = 32;
    a.kind = JOIN_PROBE; a.rcol = 0; a.ktype = Int64; a.ksize = 8;
    JoinShape b; b.robj = 2; b.lcol = 1; b.payload = PAY_ROW; b.wbits = 32;
    b.kind = JOIN_PROBE; b.rcol = 0; b.ktype = Int32; b.ksize = 4;
    s.joins = {a, b};
    s.dcol_slot[{1, 0}] = 32;
    s.dcol_slot[{2, 0}] = 33;
    KeyShape k0; k0.src = KEY_DIMROW; k0.join = 1; k0.obj = 2; k0.col = 1;
    s.keys = {k0};
    s.metrics = {"(f3)"};
    s.metric_ctype = {"int"};
    s.agg = AGG_HASH;
    s.flat = true;  // falls back to the per-row form
    out.push_back({"scan_probe", scan_checked(s)});
  }
  {  // select with a string-key probe join (Probe<char>) projecting dimension strings
    KernelShape s;
    s.fact_other = {2, 4};
    s.filter = "StrCmp(((const char*)(FCOL4+(long long)i*10)),PS(0),PI(0))==PI(1)";
    JoinShape a; a.robj = 1; a.lcol = 2; a.payload = PAY_ROW; a.wbits = 32;
    a.kind = JOIN_PROBE_STR; a.rcol = 0; a.ktype = String; a.ksize = 40;
    s.joins = {a};
    s.dcol_slot[{1, 0}] = 32;
    s.dcol_slot[{1, 1}] = 33;
    s.proj = {"((const char*)(DCOL1_1+(long long)JI1*40))", "((const char*)(DCOL1_0+(long long)JI1*40))",
              "((const char*)(FCOL2+(long long)i*40))"};
    s.proj_type = {{String, 40}, {String, 40}, {String, 40}};
    out.push_back({"select_probe_str", select_checked(s)});
  }
  {  // select with a string filter and projections, plus a row-id join
    KernelShape s;
    s.fact_i32 = {0, 2};
    s.fact_other = {1};
    s.filter = "StrCmp(((const char*)(FCOL1+(long long)i*10)),PS(0),PI(0))==PI(1)";
    JoinShape a; a.robj = 1; a.lcol = 2; a.payload = PAY_ROW; a.wbits = 32; a.lds = true;
    s.joins = {a};
    s.dcol_slot[{1, 1}] = 32;
    s.proj = {"(f0)", "((const char*)(FCOL1+(long long)i*10))",
              "(*(const int*)(DCOL1_1+(long long)JI1*4))"};
    s.proj_type = {{Int32, 4}, {String, 10}, {Int32, 4}};
    out.push_back({"select", select_checked(s)});
  }
  return out;
}

// Host-engine sources of representative shapes: scalar, dense, large dense (shared
// table), wide (packed keys), deferred maps, row-id payloads, a source-column key,
// probe joins and a select with a string-key probe.
std::vector<std::pair<std::string, std::string>> selftest_host_sources() {
  std::vector<std::pair<std::string, std::string>> out;
  {
    KernelShape s;
    s.fact_i32 = {0, 1, 2, 3};
    s.filter = "(f2)>=PI(0)&&(f2)<=PI(1)&&(f1)<PI(2)";
    JoinShape j; j.robj = 1; j.lcol = 0; j.payload = PAY_EXIST; j.wbits = 1; j.lds = true;
    s.joins = {j};
    s.metrics = {"(f3)*(f2)"};
    s.metric_ctype = {"int"};
    s.agg = AGG_SCALAR;
    out.push_back({"host_scalar", gen_host_source(s)});
  }
  {
    KernelShape s;
    s.fact_i32 = {0, 1, 2, 3, 4, 5};
    JoinShape a; a.robj = 1; a.lcol = 0; a.payload = PAY_CODE; a.wbits = 2; a.lds = true;
    JoinShape b; b.robj = 2; b.lcol = 1; b.payload = PAY_CODE; b.wbits = 8; b.defer = true;
    JoinShape c; c.robj = 3; c.lcol = 2; c.payload = PAY_ROW; c.wbits = 32;
    s.joins = {a, b, c};
    s.dcol_slot[{3, 1}] = 32;
    KeyShape k0; k0.src = KEY_DIMCODE; k0.join = 0; k0.obj = 1;
    KeyShape k1; k1.src = KEY_DIMCODE; k1.join = 1; k1.obj = 2;
    KeyShape k2; k2.src = KEY_DIMROW; k2.join = 2; k2.obj = 3; k2.col = 2;
    s.keys = {k0, k1, k2};
    s.metrics = {"(f4)-(f5)", "(*(const int*)(DCOL3_1+(long long)JI3*4))"};
    s.metric_ctype = {"int", "long long"};
    s.agg = AGG_HASH;
    out.push_back({"host_hash", gen_host_source(s)});
    s.agg = AGG_SORT;
    s.keys[2].src = KEY_FACT;
    s.keys[2].col = 3;
    out.push_back({"host_sort", gen_host_source(s)});
  }
  {
    KernelShape s;
    s.fact_i32 = {1, 3};
    s.fact_other = {0, 2};
    s.filter = "StrCmp(((const char*)(FCOL2+(long long)i*10)),PS(0),10)!=PI(1)";
    JoinShape a; a.robj = 1; a.lcol = 0; a.payload = PAY_ROW; a.wbits = 32;
    a.kind = JOIN_PROBE; a.rcol = 0; a.ktype = Int64; a.ksize = 8;
    s.joins = {a};
    s.dcol_slot[{1, 0}] = 32;
    KeyShape k0; k0.src = KEY_FACT; k0.col = 1;
    s.keys = {k0};
    s.metrics = {"(f3)"};
    s.metric_ctype = {"int"};
    s.agg = AGG_DENSE;
    out.push_back({"host_probe", gen_host_source(s)});
  }
  {
    KernelShape s;
    s.fact_other = {2, 4};
    s.filter = "StrCmp(((const char*)(FCOL4+(long long)i*10)),PS(0),PI(0))==PI(1)";
    JoinShape a; a.robj = 1; a.lcol = 2; a.payload = PAY_ROW; a.wbits = 32;
    a.kind = JOIN_PROBE_STR; a.rcol = 0; a.ktype = String; a.ksize = 40;
    s.joins = {a};
    s.dcol_slot[{1, 0}] = 32;
    s.dcol_slot[{1, 1}] = 33;
    s.proj = {"((const char*)(DCOL1_1+(long long)JI1*40))", "((const char*)(FCOL2+(long long)i*40))", "(PI(2))"};
    s.proj_type = {{String, 40}, {String, 40}, {Int64, 8}};
    out.push_back({"host_select", gen_host_source(s)});
  }
  return out;
}

}  // namespace svm

extern "C" int svm_debug_jit_selftest(char* buf, size_t cap) {
  std::string report;
  int rc = 0;
  for (auto& kv : svm::selftest_sources()) {
    try {
      std::string co = svm::jit_code_object(kv.second);
      report += kv.first + ": ok (" + std::to_string(co.size()) + " bytes)\n";
    } catch (const std::exception& e) {
      report += kv.first + ": FAILED\n" + e.what() + "\n";
      rc = -1;
    }
  }
  // the host engine's code for the same shapes, compiled by the host compiler and loaded
  for (auto& kv : svm::selftest_host_sources()) {
    try {
      (void)svm::host_task_function(kv.second);
      report += kv.first + ": host ok\n";
    } catch (const std::exception& e) {
      report += kv.first + ": host FAILED\n" + e.what() + "\n";
      rc = -1;
    }
  }
  if (buf && cap) {
    size_t n = report.size() < cap - 1 ? report.size() : cap - 1;
    memcpy(buf, report.data(), n);
    buf[n] = '\0';
  }
  return rc;
}
