// batch.cpp — batched execution: many star-join queries over one resident fact
// table fused into a shared scan, and the plain per-query fallback.
//
// The reference runs every pipeline on its own (one PipelineTest::Run per RPC,
// service/service.cu:25-63): dimension pipelines materialise filtered objects, the
// fact pipeline rebuilds host hash tables and scans its source.  A batch here:
//   1. recognises the template forms of the reference language that the shared
//      scan interprets (range / equality atoms on int columns, StrCmp equality on
//      strings, bare-column projections, metrics a, a-b, a+b, a*b);
//   2. lowers each dimension pipeline feeding a fused query into a predicate list
//      evaluated while building that query's packed map (no materialised object,
//      no compaction, no host round trip), with group codes from per-table
//      dictionaries cached across batches;
//   3. runs one persistent kernel over (tile, query) work items in tile-major
//      order, so concurrent items read the same tile through L2 / MALL;
//   4. emits every query's groups, in key order, with two batched kernels.
// Plans outside those forms run through the plain per-pipeline path unchanged.
#include "batch.h"

#include <climits>
#include <cstring>
#include <regex>
#include <condition_variable>
#include <deque>
#include <set>
#include <thread>

#include "exec.h"
#include "kernels.h"
#include "plan.h"
#include "stats.h"

namespace svm {

namespace {

// ------------------------------------------------------------ template recognition
std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && isspace((unsigned char)s[a])) a++;
  while (b > a && isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

// Strip redundant outer parentheses.
std::string unwrap(std::string s) {
  for (;;) {
    s = trim(s);
    if (s.size() < 2 || s.front() != '(' || s.back() != ')') return s;
    int depth = 0;
    bool outer = true;
    for (size_t k = 0; k < s.size(); ++k) {
      if (s[k] == '(') depth++;
      else if (s[k] == ')') depth--;
      if (depth == 0 && k + 1 < s.size()) { outer = false; break; }
    }
    if (!outer) return s;
    s = s.substr(1, s.size() - 2);
  }
}

// Split on top-level "&&"; false if the text has "||" or unbalanced parentheses.
bool split_and(const std::string& s, std::vector<std::string>& out) {
  int depth = 0;
  bool quote = false;
  size_t start = 0;
  for (size_t k = 0; k < s.size(); ++k) {
    char c = s[k];
    if (c == '"') quote = !quote;
    if (quote) continue;
    if (c == '(') depth++;
    else if (c == ')') depth--;
    else if (depth == 0 && c == '|' && k + 1 < s.size() && s[k + 1] == '|') return false;
    else if (depth == 0 && c == '&' && k + 1 < s.size() && s[k + 1] == '&') {
      out.push_back(s.substr(start, k - start));
      start = k + 2;
      k++;
    }
    if (depth < 0) return false;
  }
  if (depth != 0 || quote) return false;
  out.push_back(s.substr(start));
  return true;
}

bool plain_int(const std::string& t, long long& v) {
  // decimal literal with C meaning identical to its value (no octal, no suffix)
  std::string s = t;
  bool neg = false;
  if (!s.empty() && s[0] == '-') { neg = true; s = trim(s.substr(1)); }
  if (s.empty() || s.size() > 10) return false;
  for (char c : s) if (!isdigit((unsigned char)c)) return false;
  if (s.size() > 1 && s[0] == '0') return false;
  v = atoll(s.c_str());
  if (v > INT_MAX) return false;
  if (neg) v = -v;
  return true;
}

struct Atom {
  int obj = 0, col = 0;
  int kind = bk::DP_RANGE;
  long long lo = INT_MIN, hi = INT_MAX;
  std::string lit;
  int n = 0;
};

// "$(o.c) OP int", "int OP $(o.c)", "StrCmp($(o.c),"lit",n)==0" / "!=0"
bool parse_atom(const std::string& text, Atom& a) {
  static const std::regex re_l(R"(^\$\(\s*(\d+)\s*\.\s*(\d+)\s*\)\s*(==|<=|>=|<|>)\s*(-?\s*\d+)$)");
  static const std::regex re_r(R"(^(-?\s*\d+)\s*(==|<=|>=|<|>)\s*\$\(\s*(\d+)\s*\.\s*(\d+)\s*\)$)");
  static const std::regex re_s(
      R"RE(^StrCmp\s*\(\s*\$\(\s*(\d+)\s*\.\s*(\d+)\s*\)\s*,\s*"([^"\\]*)"\s*,\s*(\d+)\s*\)\s*(==|!=)\s*0$)RE");
  std::string s = unwrap(text);
  std::smatch m;
  std::string op;
  long long v = 0;
  if (std::regex_match(s, m, re_l)) {
    if (!plain_int(m[4].str(), v)) return false;
    a.obj = atoi(m[1].str().c_str());
    a.col = atoi(m[2].str().c_str());
    op = m[3].str();
  } else if (std::regex_match(s, m, re_r)) {
    if (!plain_int(m[1].str(), v)) return false;
    a.obj = atoi(m[3].str().c_str());
    a.col = atoi(m[4].str().c_str());
    op = m[2].str();  // mirror: v OP c  <=>  c OP' v
    if (op == "<") op = ">";
    else if (op == ">") op = "<";
    else if (op == "<=") op = ">=";
    else if (op == ">=") op = "<=";
  } else if (std::regex_match(s, m, re_s)) {
    a.obj = atoi(m[1].str().c_str());
    a.col = atoi(m[2].str().c_str());
    a.kind = m[5].str() == "==" ? bk::DP_STREQ : bk::DP_STRNE;
    a.lit = m[3].str();
    a.n = atoi(m[4].str().c_str());
    return a.lit.size() < 47;
  } else {
    return false;
  }
  if (op == "==") { a.lo = v; a.hi = v; }
  else if (op == "<") { a.hi = v - 1; }
  else if (op == "<=") { a.hi = v; }
  else if (op == ">") { a.lo = v + 1; }
  else if (op == ">=") { a.lo = v; }
  return true;
}

bool parse_conj(const std::vector<std::string>& filters, std::vector<Atom>& atoms) {
  for (auto& f : filters) {
    std::vector<std::string> parts;
    if (!split_and(f, parts)) return false;
    for (auto& p : parts) {
      Atom a;
      if (!parse_atom(p, a)) return false;
      atoms.push_back(a);
    }
  }
  return true;
}

bool parse_operand(const std::string& t, int& obj, int& col) {
  static const std::regex re(R"(^\$\(\s*(\d+)\s*\.\s*(\d+)\s*\)$)");
  std::smatch m;
  std::string s = unwrap(t);
  if (!std::regex_match(s, m, re)) return false;
  obj = atoi(m[1].str().c_str());
  col = atoi(m[2].str().c_str());
  return true;
}

bool parse_metric(const std::string& t, int& op, int& a, int& b) {
  static const std::regex re(R"(^\$\(\s*0\s*\.\s*(\d+)\s*\)\s*([-+*])\s*\$\(\s*0\s*\.\s*(\d+)\s*\)$)");
  int obj = 0;
  if (parse_operand(t, obj, a)) {
    op = bk::MOP_COL;
    b = a;
    return obj == 0;
  }
  std::smatch m;
  std::string s = unwrap(t);
  if (!std::regex_match(s, m, re)) return false;
  a = atoi(m[1].str().c_str());
  b = atoi(m[3].str().c_str());
  char c = m[2].str()[0];
  op = c == '-' ? bk::MOP_SUB : c == '+' ? bk::MOP_ADD : bk::MOP_MUL;
  return true;
}

int width_for(uint64_t maxval) {
  const int ws[] = {1, 2, 4, 8, 16};
  for (int w : ws)
    if (maxval <= ((1ull << w) - 1)) return w;
  return 32;
}

// A dimension pipeline the shared scan can absorb: filter + map of bare columns.
struct DimLowered {
  bool ok = false;
  TablePtr base;
  std::vector<Atom> atoms;
  std::vector<int> proj;  // output column q -> base column
};

DimLowered lower_dim(const PipelineInfo& info, const Plan& p, Engine& E) {
  DimLowered d;
  if (p.is_reduce || p.select_exp.empty() || !p.join_rkey.empty() || p.has_loader) return d;
  if (info.pre_pipeline_id.size() != 1) return d;
  TablePtr t = E.store.find(info.pre_pipeline_id[0]);
  if (!t) return d;
  if (!parse_conj(p.filter_exp, d.atoms)) return d;
  for (auto& a : d.atoms) {
    if (a.obj != 0 || a.col >= (int)t->cols.size()) return d;
    const Column& c = t->cols[a.col];
    if (a.kind == bk::DP_RANGE && c.type != Int32 && c.type != Int16 && c.type != Byte) return d;
    if (a.kind != bk::DP_RANGE && c.type != String) return d;
  }
  if (d.atoms.size() > (size_t)bk::kMaxDimPred) return d;
  const auto& out = p.output_info.column_list;
  for (size_t q = 0; q < p.select_exp.size(); ++q) {
    int obj, col;
    if (!parse_operand(p.select_exp[q], obj, col) || obj != 0 || col >= (int)t->cols.size()) return d;
    if (q >= out.size() || out[q].type != t->cols[col].type || out[q].val_size != t->cols[col].size) return d;
    d.proj.push_back(col);
  }
  d.base = t;
  d.ok = true;
  return d;
}

struct FactLowered {
  bool ok = false;
  TablePtr src, base;
  bk::Query q;
  std::vector<bk::MapBuild> builds;     // maps filled in later (arena offsets)
  std::vector<uint64_t> build_words;
  std::vector<std::shared_ptr<DictEntry>> dicts;  // key components in key order
  std::vector<uint32_t> strides;
  int K = 0;
  std::vector<std::string> deferred;    // dimension plan ids this query absorbs
};

struct Job {
  PipelineInfo info;
  Plan plan;
};

FactLowered lower_fact(const Job& job, const std::map<std::string, DimLowered>& dims, Engine& E,
                       hipStream_t s) {
  FactLowered f;
  const Plan& p = job.plan;
  const PipelineInfo& info = job.info;
  if (!p.is_reduce || !p.select_exp.empty() || p.has_loader || info.pre_pipeline_id.empty()) return f;
  if (p.metric_exp.size() != 1 || p.reduce_key.size() > (size_t)bk::kMaxKey) return f;
  if (p.join_rkey.empty() || p.join_rkey.size() > (size_t)bk::kMaxJoin) return f;
  TablePtr src = E.store.find(info.pre_pipeline_id[0]);
  if (!src) return f;
  TablePtr base = src->base ? src->base : src;
  if (base->cols.size() > (size_t)bk::kMaxCols) return f;
  const auto& out = p.output_info.column_list;
  const size_t K = p.reduce_key.size();
  if (out.size() < K + 1) return f;
  for (size_t q = 0; q < K; ++q)
    if (out[q].type != Int32) return f;
  if (out[K].type != Int32 && out[K].type != Int64) return f;
  // ref objects: deferred dimension pipelines of this batch or resident tables
  std::vector<DimLowered> refs;
  for (size_t k = 1; k < info.pre_pipeline_id.size(); ++k) {
    auto it = dims.find(info.pre_pipeline_id[k]);
    if (it != dims.end()) {
      if (!it->second.ok) return f;
      refs.push_back(it->second);
      f.deferred.push_back(info.pre_pipeline_id[k]);
      continue;
    }
    TablePtr t = E.store.find(info.pre_pipeline_id[k]);
    if (!t) return f;
    DimLowered d;
    d.ok = true;
    d.base = t;
    for (size_t c = 0; c < t->cols.size(); ++c) d.proj.push_back((int)c);
    refs.push_back(d);
  }
  auto src_int = [&](int col) { return col >= 0 && col < (int)src->cols.size() && src->cols[col].type == Int32; };
  bk::Query& q = f.q;
  memset(&q, 0, sizeof(q));
  q.row0 = (long long)src->row0;
  q.row1 = (long long)(src->row0 + src->rows);
  // fact predicates: intersect ranges per column
  std::vector<Atom> atoms;
  if (!parse_conj(p.filter_exp, atoms)) return f;
  std::map<int, std::pair<long long, long long>> ranges;
  for (auto& a : atoms) {
    if (a.obj != 0 || a.kind != bk::DP_RANGE || !src_int(a.col)) return f;
    auto& r = ranges.emplace(a.col, std::make_pair((long long)INT_MIN, (long long)INT_MAX)).first->second;
    r.first = std::max(r.first, a.lo);
    r.second = std::min(r.second, a.hi);
  }
  if (ranges.size() > (size_t)bk::kMaxPred) return f;
  for (auto& kv : ranges) {
    q.pcol[q.npred] = kv.first;
    q.plo[q.npred] = (int)std::max<long long>(kv.second.first, INT_MIN);
    q.phi[q.npred] = (int)std::min<long long>(kv.second.second, INT_MAX);
    if (kv.second.first > kv.second.second) { q.plo[q.npred] = 1; q.phi[q.npred] = 0; }
    q.npred++;
  }
  // metric
  if (!parse_metric(p.metric_exp[0], q.mop, q.ma, q.mb) || !src_int(q.ma) || !src_int(q.mb)) return f;
  // keys: one component per dimension, read through that dimension's join
  std::vector<int> key_obj(K), key_base(K);
  for (size_t k = 0; k < K; ++k) {
    KeyIndex ki = p.reduce_key[k];
    if (ki.obj < 1 || ki.obj > (int)refs.size()) return f;
    const DimLowered& d = refs[ki.obj - 1];
    if (ki.col < 0 || ki.col >= (int)d.proj.size()) return f;
    key_obj[k] = ki.obj;
    key_base[k] = d.proj[ki.col];
    int t = d.base->cols[key_base[k]].type;
    if (t != Int32 && t != String && t != Int16 && t != Byte) return f;
    for (size_t k2 = 0; k2 < k; ++k2)
      if (key_obj[k2] == ki.obj) return f;
  }
  uint64_t G = 1;
  f.dicts.resize(K);
  f.strides.resize(K);
  for (int k = (int)K - 1; k >= 0; --k) {
    f.dicts[k] = table_dict(refs[key_obj[k] - 1].base, key_base[k], s);
    f.strides[k] = (uint32_t)G;
    G *= f.dicts[k]->n;
    if (G > (1ull << 22)) return f;
  }
  f.K = (int)K;
  q.G = (unsigned)(K ? G : 1);
  q.agg = K == 0 ? bk::BAGG_SCALAR : G <= (uint64_t)bk::kDenseG ? bk::BAGG_DENSE : bk::BAGG_HASH;
  // joins
  std::set<int> seen;
  for (size_t j = 0; j < p.join_rkey.size(); ++j) {
    KeyIndex L = p.join_lkey[j], R = p.join_rkey[j];
    if (L.obj != 0 || !src_int(L.col) || R.obj < 1 || R.obj > (int)refs.size()) return f;
    if (!seen.insert(R.obj).second) return f;
    const DimLowered& d = refs[R.obj - 1];
    if (R.col < 0 || R.col >= (int)d.proj.size()) return f;
    int rcol = d.proj[R.col];
    if (d.base->cols[rcol].type != Int32) return f;
    if (d.base->max_tuple_count && d.base->rows > d.base->max_tuple_count) return f;
    KeyEntry ke = table_key(d.base, rcol, s);
    if (!ke.unique) return f;
    bk::MapBuild b;
    memset(&b, 0, sizeof(b));
    b.keys = (const int32_t*)d.base->col_ptr(rcol);
    b.n = d.base->rows;
    b.kmin = ke.kmin;
    uint64_t range = d.base->rows ? (uint64_t)((int64_t)ke.kmax - (int64_t)ke.kmin + 1) : 0;
    if (range > (1ull << 28)) return f;
    b.range = (unsigned)range;
    for (auto& a : d.atoms) {
      bk::DimPred& dp = b.pred[b.npred++];
      dp.col = d.base->col_ptr(a.col);
      dp.kind = a.kind;
      dp.type = d.base->cols[a.col].type;
      dp.size = (int)d.base->cols[a.col].size;
      dp.lo = (int)std::max<long long>(a.lo, INT_MIN);
      dp.hi = (int)std::min<long long>(a.hi, INT_MAX);
      if (a.lo > a.hi) { dp.lo = 1; dp.hi = 0; }
      dp.n = a.n;
      memset(dp.lit, 0, sizeof(dp.lit));
      memcpy(dp.lit, a.lit.data(), a.lit.size());
    }
    int comp = -1;
    for (size_t k = 0; k < K; ++k)
      if (key_obj[k] == R.obj) comp = (int)k;
    b.codes = comp >= 0 ? f.dicts[comp]->codes.as<int32_t>() : nullptr;
    b.wbits = comp >= 0 ? width_for(f.dicts[comp]->n) : 1;
    q.jcol[q.njoin] = L.col;
    q.jkmin[q.njoin] = b.kmin;
    q.jrange[q.njoin] = b.range;
    q.jw[q.njoin] = b.wbits;
    q.jcode[q.njoin] = comp >= 0 ? 1 : 0;
    q.jstride[q.njoin] = comp >= 0 ? f.strides[comp] : 0;
    q.njoin++;
    uint64_t per = 32 / (uint64_t)b.wbits;
    f.build_words.push_back(((uint64_t)b.range + per - 1) / per);
    f.builds.push_back(b);
  }
  // every key component's dimension must be joined
  for (size_t k = 0; k < K; ++k)
    if (!seen.count(key_obj[k])) return f;
  f.src = src;
  f.base = base;
  f.ok = true;
  return f;
}

}  // namespace

// Device buffers reused across batches (grown on demand, never shrunk), so a
// steady stream of batches allocates nothing.
struct Scratch {
  std::vector<DevBuf> bufs;
  void* host = nullptr;  // pinned staging for descriptors
  size_t host_bytes = 0;
  void* get(size_t slot, uint64_t bytes) {
    if (bufs.size() <= slot) bufs.resize(slot + 1);
    if (bufs[slot].bytes < bytes) bufs[slot] = DevBuf(bytes + bytes / 4 + 256);
    return bufs[slot].p;
  }
  char* pinned(size_t bytes) {
    if (host_bytes < bytes) {
      if (host) (void)hipHostFree(host);
      host = nullptr;
      size_t b = bytes + bytes / 4 + 4096;
      SVM_HIP(hipHostMalloc(&host, b, hipHostMallocDefault));
      host_bytes = b;
    }
    return (char*)host;
  }
  ~Scratch() {
    if (host) (void)hipHostFree(host);
  }
};

// HBM a fused query holds while it runs: its maps and group tables.
static uint64_t fused_bytes(const FactLowered& f) {
  uint64_t b = 4096 + (uint64_t)f.q.G * 8 + ((uint64_t)f.q.G + 31) / 32 * 4;
  for (uint64_t w : f.build_words) b += w * 4 + 16;
  return b;
}

// Fused execution of lowered queries that share one base fact table.
static void run_fused(std::vector<FactLowered*>& qs, std::vector<std::unique_ptr<svm_result>*>& outs,
                      Engine& E, hipStream_t s, Scratch& sc) {
  if (qs.empty()) return;
  TablePtr base = qs[0]->base;
  const int nq = (int)qs.size();
  const long long tile = 65536;
  // arenas
  uint64_t map_words = 0, sum_words = 0, flag_words = 0, nbuild = 0, max_rows = 0;
  for (auto* f : qs) {
    for (size_t j = 0; j < f->builds.size(); ++j) {
      map_words += (f->build_words[j] + 3) & ~3ull;
      max_rows = std::max<uint64_t>(max_rows, f->builds[j].n);
    }
    nbuild += f->builds.size();
    sum_words += f->q.G;
    flag_words += (f->q.G + 31) / 32;
  }
  hipEvent_t e0, e1;
  SVM_HIP(hipEventCreate(&e0));
  SVM_HIP(hipEventCreate(&e1));
  SVM_HIP(hipEventRecord(e0, s));
  unsigned* maps = (unsigned*)sc.get(0, map_words * 4 + 16);
  unsigned long long* sums = (unsigned long long*)sc.get(1, sum_words * 8 + 8);
  unsigned* flags = (unsigned*)sc.get(2, flag_words * 4 + 4);
  unsigned* counter = (unsigned*)sc.get(3, 16);
  k::fill_u32(maps, map_words + 4, 0, s);
  k::fill_u32(sums, sum_words * 2 + 2, 0, s);
  k::fill_u32(flags, flag_words + 1, 0, s);
  k::fill_u32(counter, 4, 0, s);
  std::vector<bk::MapBuild> builds;
  std::vector<bk::Query> qd;
  std::vector<bk::Emit> em;
  uint64_t mo = 0, so = 0, fo = 0, blocks = 0;
  for (auto* f : qs) {
    bk::Query q = f->q;
    for (size_t j = 0; j < f->builds.size(); ++j) {
      bk::MapBuild b = f->builds[j];
      b.map = maps + mo;
      q.jmap[j] = b.map;
      mo += (f->build_words[j] + 3) & ~3ull;
      builds.push_back(b);
    }
    q.sums = sums + so;
    q.flags = flags + fo;
    bk::Emit e;
    memset(&e, 0, sizeof(e));
    e.flags = q.flags;
    e.sums = q.sums;
    e.G = q.G;
    e.K = f->K;
    for (int k = 0; k < f->K; ++k) {
      e.n[k] = f->dicts[k]->n;
      e.stride[k] = f->strides[k];
      e.dict[k] = f->dicts[k]->dict.as<int32_t>();
    }
    e.block0 = blocks;
    e.nblocks = ((q.G + 31) / 32 + 255) / 256;
    blocks += e.nblocks;
    so += q.G;
    fo += (q.G + 31) / 32;
    qd.push_back(q);
    em.push_back(e);
  }
  // work items: tile-major so concurrent workgroups share a tile through the caches
  std::vector<bk::Item> items;
  long long tmax = 0;
  for (auto& q : qd) tmax = std::max(tmax, (q.row1 + tile - 1) / tile);
  for (long long t = 0; t < tmax; ++t)
    for (int i = 0; i < nq; ++i)
      if (qd[i].row0 < (t + 1) * tile && qd[i].row1 > t * tile && qd[i].row1 > qd[i].row0)
        items.push_back({(int)t, i});
  // descriptors in one upload
  size_t ob = 0, oq = ob + builds.size() * sizeof(bk::MapBuild);
  oq = (oq + 255) & ~255ull;
  size_t oi = oq + qd.size() * sizeof(bk::Query);
  oi = (oi + 255) & ~255ull;
  size_t oe = oi + items.size() * sizeof(bk::Item);
  oe = (oe + 255) & ~255ull;
  size_t total = oe + em.size() * sizeof(bk::Emit);
  char* host = sc.pinned(total + 256);
  memcpy(host + ob, builds.data(), builds.size() * sizeof(bk::MapBuild));
  memcpy(host + oq, qd.data(), qd.size() * sizeof(bk::Query));
  memcpy(host + oi, items.data(), items.size() * sizeof(bk::Item));
  memcpy(host + oe, em.data(), em.size() * sizeof(bk::Emit));
  char* dd = (char*)sc.get(4, total + 256);
  SVM_HIP(hipMemcpyAsync(dd, host, total, hipMemcpyHostToDevice, s));
  hipEvent_t eb, es;
  SVM_HIP(hipEventCreate(&eb));
  SVM_HIP(hipEventCreate(&es));
  bk::launch_map_build((const bk::MapBuild*)(dd + ob), (int)builds.size(), max_rows, s);
  SVM_HIP(hipEventRecord(eb, s));
  const void* cols[bk::kMaxCols] = {nullptr};
  for (size_t c = 0; c < base->cols.size() && c < (size_t)bk::kMaxCols; ++c) cols[c] = base->cols[c].dev;
  if (!items.empty())
    bk::launch_shared_scan((const bk::Query*)(dd + oq), (const bk::Item*)(dd + oi), (unsigned)items.size(),
                           counter, cols, tile, E.num_cus * 4, s);
  SVM_HIP(hipEventRecord(es, s));
  unsigned long long* counts = (unsigned long long*)sc.get(5, blocks * 8 + 8);
  unsigned long long* offsets = (unsigned long long*)sc.get(6, blocks * 8 + 8);
  unsigned long long* totals = (unsigned long long*)sc.get(7, nq * 8 + 8);
  bk::launch_emit_count((const bk::Emit*)(dd + oe), nq, blocks, counts, offsets, totals, s);
  std::vector<unsigned long long> tot(nq);
  SVM_HIP(hipMemcpyAsync(tot.data(), totals, nq * 8, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  uint64_t rows = 0;
  for (int i = 0; i < nq; ++i) {
    em[i].out0 = rows;
    rows += tot[i];
  }
  memcpy(host + oe, em.data(), em.size() * sizeof(bk::Emit));
  SVM_HIP(hipMemcpyAsync(dd + oe, host + oe, em.size() * sizeof(bk::Emit), hipMemcpyHostToDevice, s));
  int32_t* keys = (int32_t*)sc.get(8, rows * 4 * bk::kMaxKey + 16);
  int64_t* vals = (int64_t*)sc.get(9, rows * 8 + 8);
  bk::launch_emit_rows((const bk::Emit*)(dd + oe), nq, blocks, offsets, keys, vals, rows, s);
  std::vector<int32_t> hk(rows * bk::kMaxKey + 1);
  std::vector<int64_t> hv(rows + 1);
  if (rows) {
    SVM_HIP(hipMemcpyAsync(hk.data(), keys, rows * 4 * bk::kMaxKey, hipMemcpyDeviceToHost, s));
    SVM_HIP(hipMemcpyAsync(hv.data(), vals, rows * 8, hipMemcpyDeviceToHost, s));
  }
  SVM_HIP(hipEventRecord(e1, s));
  SVM_HIP(hipStreamSynchronize(s));
  float ms = 0, ms_build = 0, ms_scan = 0;
  SVM_HIP(hipEventElapsedTime(&ms, e0, e1));
  SVM_HIP(hipEventElapsedTime(&ms_build, e0, eb));
  SVM_HIP(hipEventElapsedTime(&ms_scan, eb, es));
  if (E.cfg.verbose) {
    uint64_t slice_rows = 0, dim_rows = 0;
    for (auto& q : qd) slice_rows += (uint64_t)(q.row1 - q.row0);
    for (auto& b : builds) dim_rows += b.n;
    fprintf(stderr, "[svm] batch of %d: gpu %.3f ms (map builds %.3f over %llu dim rows, shared scan %.3f over "
                    "%llu slice rows / %zu items, emit %.3f)\n",
            nq, ms, ms_build, (unsigned long long)dim_rows, ms_scan, (unsigned long long)slice_rows,
            items.size(), ms - ms_build - ms_scan);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(eb);
  (void)hipEventDestroy(es);
  for (int i = 0; i < nq; ++i) {
    auto r = std::make_unique<svm_result>();
    const int K = qs[i]->K;
    r->kind = K == 0 ? 2 : 1;
    r->rows = (int64_t)tot[i];
    for (int k = 0; k < K; ++k) {
      r->types.push_back({Int32, 4});
      std::vector<uint8_t> col(tot[i] * 4);
      if (tot[i]) memcpy(col.data(), hk.data() + (size_t)k * rows + em[i].out0, tot[i] * 4);
      r->cols.push_back(std::move(col));
    }
    r->types.push_back({Int64, 8});
    std::vector<uint8_t> col(tot[i] * 8);
    if (tot[i]) memcpy(col.data(), hv.data() + em[i].out0, tot[i] * 8);
    r->cols.push_back(std::move(col));
    r->scan_ms = ms;  // the whole batch completes together
    *outs[i] = std::move(r);
  }
}

int batch_run(int n, const uint8_t* const* plans, const size_t* lens, svm_result** outs) {
  if (n < 0 || (n && (!plans || !lens))) fail(SVM_ERR_INVALID, "bad batch arguments");
  Engine& E = engine();
  std::lock_guard<std::mutex> g(E.run_mu);
  hipStream_t s = E.stream;
  const auto t0 = std::chrono::steady_clock::now();
  auto host_ms = [&] {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  };
  std::vector<Job> jobs(n);
  for (int i = 0; i < n; ++i) {
    jobs[i].info = decode_pipeline_info(plans[i], lens[i]);
    jobs[i].plan = decode_plan(jobs[i].info);
  }
  // dimension candidates, by pipeline id
  std::map<std::string, DimLowered> dims;
  for (auto& j : jobs) {
    DimLowered d = lower_dim(j.info, j.plan, E);
    if (d.ok) dims[j.info.id] = d;
  }
  std::vector<FactLowered> facts(n);
  std::vector<bool> fused(n, false);
  for (int i = 0; i < n; ++i) {
    if (!jobs[i].plan.is_reduce) continue;
    facts[i] = lower_fact(jobs[i], dims, E, s);
    fused[i] = facts[i].ok;
  }
  // a dimension pipeline is absorbed only if no plain plan reads its output
  std::set<std::string> needed;
  for (int i = 0; i < n; ++i) {
    if (fused[i]) continue;
    for (size_t k = 0; k < jobs[i].info.pre_pipeline_id.size(); ++k) needed.insert(jobs[i].info.pre_pipeline_id[k]);
  }
  std::set<std::string> absorbed;
  for (int i = 0; i < n; ++i)
    if (fused[i])
      for (auto& id : facts[i].deferred) absorbed.insert(id);
  for (auto& id : needed) absorbed.erase(id);
  for (int i = 0; i < n; ++i)
    if (fused[i])
      for (auto& id : facts[i].deferred)
        if (!absorbed.count(id)) fused[i] = false;  // its dimension is materialised instead
  std::vector<std::unique_ptr<svm_result>> res(n);
  const double t_lowered = host_ms();
  // plain plans in submission order (absorbed dimension pipelines are skipped)
  for (int i = 0; i < n; ++i) {
    if (fused[i] || absorbed.count(jobs[i].info.id)) continue;
    RunOptions opt;
    opt.want_host = outs != nullptr;
    res[i] = run_pipeline(jobs[i].info, opt);
  }
  const double t_plain = host_ms();
  // fused queries, one shared scan per base fact table
  std::map<const Table*, std::vector<int>> groups;
  for (int i = 0; i < n; ++i)
    if (fused[i]) groups[facts[i].base.get()].push_back(i);
  for (auto& gq : groups) {
    std::vector<FactLowered*> fq;
    std::vector<std::unique_ptr<svm_result>*> fo;
    for (int i : gq.second) {
      fq.push_back(&facts[i]);
      fo.push_back(&res[i]);
    }
    static Scratch* scratch = new Scratch();  // reused by every batch (guarded by run_mu)
    run_fused(fq, fo, E, s, *scratch);
  }
  int nfused = 0;
  for (int i = 0; i < n; ++i) nfused += fused[i] ? 1 : 0;
  if (outs)
    for (int i = 0; i < n; ++i) outs[i] = res[i] ? res[i].release() : nullptr;
  if (E.cfg.verbose)
    fprintf(stderr, "[svm] batch_run %d plans, %d fused: host lowering %.3f ms, plain plans %.3f ms, total %.3f ms\n",
            n, nfused, t_lowered, t_plain - t_lowered, host_ms());
  return nfused;
}

// ------------------------------------------------------------------ serving
namespace {

struct Ticket {
  uint64_t id = 0;
  std::vector<Job> jobs;  // dimension pipelines, then the fact pipeline
  FactLowered fact;
  bool fused = false;
  uint64_t need = 0;
  bool done = false;
  int status = SVM_OK;
  std::string error;
  std::unique_ptr<svm_result> result;
};

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Ticket>> pending;
  std::map<uint64_t, std::shared_ptr<Ticket>> tickets;
  uint64_t next_id = 1, bytes_in_use = 0, budget = 0, running = 0, completed = 0;
  int max_batch = 64, mode = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  hipStream_t submit_stream = nullptr;
};

std::unique_ptr<Server> g_server;
std::mutex g_server_mu;
constexpr uint64_t kPlainNeed = 64ull << 20;  // HBM assumed for a query on the plain path
constexpr uint64_t kSharedScanMaxRows = 64ull << 20;  // larger slices take the JIT-specialised path

void run_plain_query(Ticket& t, hipStream_t s) {
  std::unique_ptr<svm_result> last;
  for (auto& j : t.jobs) {
    RunOptions opt;
    opt.stream = s;
    opt.want_host = true;
    last = run_pipeline(j.info, opt);
  }
  t.result = std::move(last);
}

void worker_main(Server* S, int index) {
  Engine& E = engine();
  hipStream_t s = nullptr;
  SVM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  Scratch sc;
  (void)index;
  for (;;) {
    std::vector<std::shared_ptr<Ticket>> batch;
    uint64_t need = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] {
        if (S->stop) return true;
        return !S->pending.empty() && S->bytes_in_use + S->pending.front()->need <= S->budget;
      });
      if (S->stop && S->pending.empty()) break;
      if (S->stop) break;
      // admit in arrival order while memory and the batch size allow
      const int cap = S->mode == 1 ? 1 : S->max_batch;
      while (!S->pending.empty() && (int)batch.size() < cap &&
             S->bytes_in_use + need + S->pending.front()->need <= S->budget) {
        auto t = S->pending.front();
        if (S->mode == 0 && !batch.empty() && (!t->fused || !batch[0]->fused ||
                                               t->fact.base != batch[0]->fact.base))
          break;  // a batch is one shared scan over one fact table
        need += t->need;
        batch.push_back(t);
        S->pending.pop_front();
        if (!t->fused) break;
      }
      S->bytes_in_use += need;
      S->running += batch.size();
    }
    try {
      if (batch.size() == 1 && (!batch[0]->fused || S->mode == 1)) {
        run_plain_query(*batch[0], s);
      } else {
        std::vector<FactLowered*> fq;
        std::vector<std::unique_ptr<svm_result>*> fo;
        for (auto& t : batch) {
          fq.push_back(&t->fact);
          fo.push_back(&t->result);
        }
        run_fused(fq, fo, E, s, sc);
      }
    } catch (const Error& e) {
      for (auto& t : batch) { t->status = e.code; t->error = e.what(); }
    } catch (const std::exception& e) {
      for (auto& t : batch) { t->status = SVM_ERR_DEVICE; t->error = e.what(); }
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->bytes_in_use -= need;
      S->running -= batch.size();
      S->completed += batch.size();
      for (auto& t : batch) t->done = true;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
}

}  // namespace

int server_start(int workers, uint64_t budget, int max_batch, int mode) {
  std::lock_guard<std::mutex> g(g_server_mu);
  Engine& E = engine();
  if (g_server) fail(SVM_ERR_STATE, "server already running");
  if (workers <= 0 || workers > 16) fail(SVM_ERR_INVALID, "workers must be 1..16");
  if (mode != 0 && mode != 1) fail(SVM_ERR_INVALID, "mode is 0 (continuous batching) or 1 (plain)");
  auto S = std::make_unique<Server>();
  S->budget = budget ? budget : E.cfg.hbm_budget;
  S->max_batch = max_batch > 0 ? max_batch : 64;
  S->mode = mode;
  SVM_HIP(hipStreamCreateWithFlags(&S->submit_stream, hipStreamNonBlocking));
  Server* raw = S.get();
  for (int w = 0; w < workers; ++w) S->workers.emplace_back(worker_main, raw, w);
  g_server = std::move(S);
  return SVM_OK;
}

int server_submit(const uint8_t* const* plans, const size_t* lens, int n, uint64_t* ticket) {
  Server* S = g_server.get();
  if (!S) fail(SVM_ERR_STATE, "server not running");
  if (n <= 0 || !plans || !lens) fail(SVM_ERR_INVALID, "a query needs at least its fact pipeline");
  Engine& E = engine();
  auto t = std::make_shared<Ticket>();
  t->jobs.resize(n);
  for (int i = 0; i < n; ++i) {
    t->jobs[i].info = decode_pipeline_info(plans[i], lens[i]);
    t->jobs[i].plan = decode_plan(t->jobs[i].info);
  }
  if (S->mode == 0) {
    std::map<std::string, DimLowered> dims;
    for (int i = 0; i + 1 < n; ++i) {
      DimLowered d = lower_dim(t->jobs[i].info, t->jobs[i].plan, E);
      if (d.ok) dims[t->jobs[i].info.id] = d;
    }
    t->fact = lower_fact(t->jobs[n - 1], dims, E, S->submit_stream);
    // every dimension pipeline of the query must be absorbed by the fused fact pipeline
    t->fused = t->fact.ok && t->fact.deferred.size() == (size_t)(n - 1);
    // a query over a large slice runs faster on its own specialised kernel (the
    // shared scan interprets descriptors); it still overlaps with batches on
    // another worker
    if (t->fused && (uint64_t)(t->fact.q.row1 - t->fact.q.row0) > kSharedScanMaxRows) t->fused = false;
  }
  t->need = t->fused ? fused_bytes(t->fact) : kPlainNeed;
  if (t->need > S->budget) fail(SVM_ERR_BUSY, "query needs more HBM than the serving budget");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) fail(SVM_ERR_STATE, "server stopping");
    t->id = S->next_id++;
    S->tickets[t->id] = t;
    S->pending.push_back(t);
    if (ticket) *ticket = t->id;
  }
  S->cv_work.notify_one();
  return SVM_OK;
}

int server_wait(uint64_t ticket, int timeout_ms, svm_result** out) {
  Server* S = g_server.get();
  if (!S) fail(SVM_ERR_STATE, "server not running");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) fail(SVM_ERR_NOT_FOUND, "unknown ticket " + std::to_string(ticket));
  auto t = it->second;
  auto ready = [&] { return t->done; };
  if (timeout_ms < 0) S->cv_done.wait(lk, ready);
  else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
    return SVM_ERR_BUSY;
  S->tickets.erase(ticket);
  if (t->status != SVM_OK) fail(t->status, t->error);
  if (out) *out = t->result.release();
  return SVM_OK;
}

int server_stats(uint64_t* queued, uint64_t* running, uint64_t* done, uint64_t* bytes) {
  Server* S = g_server.get();
  if (!S) fail(SVM_ERR_STATE, "server not running");
  std::lock_guard<std::mutex> lk(S->mu);
  if (queued) *queued = S->pending.size();
  if (running) *running = S->running;
  if (done) *done = S->completed;
  if (bytes) *bytes = S->bytes_in_use;
  return SVM_OK;
}

int server_stop() {
  std::lock_guard<std::mutex> g(g_server_mu);
  if (!g_server) return SVM_OK;
  {
    std::lock_guard<std::mutex> lk(g_server->mu);
    g_server->stop = true;
  }
  g_server->cv_work.notify_all();
  for (auto& w : g_server->workers) w.join();
  {
    // queries never admitted fail with SVM_ERR_STATE
    std::lock_guard<std::mutex> lk(g_server->mu);
    for (auto& t : g_server->pending) {
      t->status = SVM_ERR_STATE;
      t->error = "server stopped before the query was admitted";
      t->done = true;
    }
    g_server->pending.clear();
  }
  g_server->cv_done.notify_all();
  (void)hipStreamDestroy(g_server->submit_stream);
  g_server.reset();
  return SVM_OK;
}

}  // namespace svm
