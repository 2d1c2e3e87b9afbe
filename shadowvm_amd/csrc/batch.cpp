// batch.cpp — batched execution: many star-join queries over one resident fact
// table fused into a shared scan, and the plain per-query fallback.
//
// The reference runs every pipeline on its own (one PipelineTest::Run per RPC,
// service/service.cu:25-63): dimension pipelines materialise filtered objects, the
// fact pipeline rebuilds host hash tables and scans its source.  A batch here:
//   1. recognises the template forms of the reference language that the shared
//      scan interprets (range / equality atoms on int columns, StrCmp equality on
//      strings, bare-column projections, metrics a, a-b, a+b, a*b);
//   2. lowers each dimension pipeline feeding a fused query into a list of atoms
//      (no materialised object, no compaction, no host round trip);
//   3. groups the batch's joins into slots (dimension table, key column, fact
//      column); each slot's dimension rows carry a profile — the packed dictionary
//      codes of the columns the batch's queries filter or group on, deduplicated
//      and cached with the table — and one small kernel evaluates every query's
//      atoms on every profile into a 64-bit mask;
//   4. runs one persistent kernel over tiles of the fact table: each row is read
//      once, looks up one profile per slot and aggregates every (row, query) pair
//      that passes into an LDS table keyed by (query, group);
//   5. emits every query's groups, in key order, with two batched kernels.
// Plans outside those forms run through the plain per-pipeline path unchanged.
#include "batch.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>

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

// A cursor over the text of an atom, for the fixed grammar of the fusable forms
// (white space as C isspace, digits 0-9).
struct Scan {
  const char* p;
  const char* e;
  explicit Scan(const std::string& s) : p(s.data()), e(s.data() + s.size()) {}
  void ws() { while (p < e && isspace((unsigned char)*p)) ++p; }
  bool lit(const char* t) {
    const size_t n = strlen(t);
    if ((size_t)(e - p) < n || memcmp(p, t, n) != 0) return false;
    p += n;
    return true;
  }
  bool digits(std::string& out) {
    const char* b = p;
    while (p < e && *p >= '0' && *p <= '9') ++p;
    if (p == b) return false;
    out.assign(b, p);
    return true;
  }
  bool end() const { return p == e; }
  // $( o . c ), with white space allowed inside; the object's digits are kept as text
  bool operand(std::string& obj, int& col) {
    std::string c;
    if (!lit("$(")) return false;
    ws();
    if (!digits(obj)) return false;
    ws();
    if (!lit(".")) return false;
    ws();
    if (!digits(c)) return false;
    ws();
    if (!lit(")")) return false;
    col = atoi(c.c_str());
    return true;
  }
  // == <= >= < > (tried in that order)
  bool cmp(std::string& op) {
    for (const char* o : {"==", "<=", ">=", "<", ">"})
      if (lit(o)) {
        op = o;
        return true;
      }
    return false;
  }
  // an optional minus, white space, digits
  bool int_text(std::string& t) {
    const char* b = p;
    lit("-");
    ws();
    std::string d;
    if (!digits(d)) return false;
    t.assign(b, p);
    return true;
  }
};

// "$(o.c) OP int", "int OP $(o.c)", "StrCmp($(o.c),"lit",n)==0" / "!=0"
bool parse_atom(const std::string& text, Atom& a) {
  const std::string s = unwrap(text);
  std::string op, num, obj;
  long long v = 0;
  Scan l(s);
  if (l.operand(obj, a.col) && (l.ws(), l.cmp(op)) && (l.ws(), l.int_text(num)) && l.end()) {
    if (!plain_int(num, v)) return false;
    a.obj = atoi(obj.c_str());
  } else if (Scan r(s); r.int_text(num) && (r.ws(), r.cmp(op)) && (r.ws(), r.operand(obj, a.col)) && r.end()) {
    if (!plain_int(num, v)) return false;
    a.obj = atoi(obj.c_str());
    // mirror: v OP c  <=>  c OP' v
    if (op == "<") op = ">";
    else if (op == ">") op = "<";
    else if (op == "<=") op = ">=";
    else if (op == ">=") op = "<=";
  } else {
    Scan t(s);
    std::string n, lit;
    if (!t.lit("StrCmp")) return false;
    t.ws();
    if (!t.lit("(")) return false;
    t.ws();
    if (!t.operand(obj, a.col)) return false;
    t.ws();
    if (!t.lit(",")) return false;
    t.ws();
    if (!t.lit("\"")) return false;
    const char* b = t.p;
    while (t.p < t.e && *t.p != '"' && *t.p != '\\') ++t.p;
    lit.assign(b, t.p);
    if (!t.lit("\"")) return false;
    t.ws();
    if (!t.lit(",")) return false;
    t.ws();
    if (!t.digits(n)) return false;
    t.ws();
    if (!t.lit(")")) return false;
    t.ws();
    if (!t.cmp(op) && !t.lit("!=")) return false;
    const bool ne = op.empty();
    if (!ne && op != "==") return false;
    t.ws();
    if (!t.lit("0") || !t.end()) return false;
    a.obj = atoi(obj.c_str());
    a.kind = ne ? bk::DP_STRNE : bk::DP_STREQ;
    a.lit = lit;
    a.n = atoi(n.c_str());
    return a.lit.size() < 47;
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
  const std::string s = unwrap(t);
  Scan c(s);
  std::string o;
  if (!c.operand(o, col) || !c.end()) return false;
  obj = atoi(o.c_str());
  return true;
}

// "$(0.a)" or "$(0.a) OP $(0.b)" with OP one of - + *
bool parse_metric(const std::string& t, int& op, int& a, int& b) {
  int obj = 0;
  if (parse_operand(t, obj, a)) {
    op = bk::MOP_COL;
    b = a;
    return obj == 0;
  }
  const std::string s = unwrap(t);
  Scan c(s);
  std::string o1, o2;
  if (!c.operand(o1, a) || o1 != "0") return false;
  c.ws();
  if (c.p == c.e || (*c.p != '-' && *c.p != '+' && *c.p != '*')) return false;
  const char x = *c.p++;
  c.ws();
  if (!c.operand(o2, b) || o2 != "0" || !c.end()) return false;
  op = x == '-' ? bk::MOP_SUB : x == '+' ? bk::MOP_ADD : bk::MOP_MUL;
  return true;
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
  // a pipeline that asks for the host engine (isGPU false) is not absorbed into a shared scan
  if (!info.isGPU || p.is_reduce || p.select_exp.empty() || !p.join_rkey.empty() || p.has_loader) return d;
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

// ------------------------------------------------------------------ profiles
// A string column's dictionary holds Hash31 values (the reference's group key of
// a string); a dimension atom may be evaluated on a code only when every row with
// that code has the same bytes.  Checked once per dictionary and cached.
struct StrRep {
  std::weak_ptr<DictEntry> owner;
  DevBuf rep;  // first row of each code
  bool exact = false;
};
std::mutex g_prof_mu;
std::map<const DictEntry*, std::shared_ptr<StrRep>> g_strrep;

std::shared_ptr<StrRep> string_rep(const TablePtr& t, int col, const std::shared_ptr<DictEntry>& d, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  auto it = g_strrep.find(d.get());
  if (it != g_strrep.end() && it->second->owner.lock() == d) return it->second;
  auto r = std::make_shared<StrRep>();
  r->owner = d;
  r->rep = DevBuf((uint64_t)d->n * 4 + 4);
  if (t->rows && d->n) {
    k::first_rows(d->codes.as<int32_t>(), t->rows, 0, d->n, r->rep.as<int32_t>(), s);
    DevBuf flag(4);
    k::fill_u32(flag.p, 1, 0, s);
    bk::launch_string_exact((const char*)t->col_ptr(col), (int)t->cols[col].size, t->rows, d->codes.as<int32_t>(),
                            r->rep.as<int32_t>(), flag.as<unsigned>(), s);
    unsigned h = 1;
    SVM_HIP(hipMemcpyAsync(&h, flag.p, 4, hipMemcpyDeviceToHost, s));
    SVM_HIP(hipStreamSynchronize(s));
    r->exact = h == 0;
  } else {
    r->exact = true;
  }
  g_strrep[d.get()] = r;
  return r;
}

int bits_for(uint32_t n) {
  int b = 0;
  while (b < 31 && (1ull << b) < n) b++;
  return b;
}

// Profiles of one (dimension table, key column, attribute set): every dimension
// row's attribute codes packed into `raw`, deduplicated into P sorted profiles, and
// the profile of each key offset.
struct Profile {
  std::weak_ptr<Table> owner;
  uint64_t version = 0;  // owner's data_version() when built
  std::vector<int> attrs, shift, bits;
  std::vector<std::shared_ptr<DictEntry>> dicts;
  uint32_t P = 0;
  bool ok = false;
  int32_t kmin = 0;
  uint32_t range = 0;
  DevBuf praw;  // P packed codes, ascending
  DevBuf pkey;  // per key offset: uint8 when narrow (P < 255), else uint16; all ones = no row
  bool narrow = false;
};
using ProfileKey = std::tuple<const Table*, int, std::vector<int>>;
std::map<ProfileKey, std::shared_ptr<Profile>> g_profiles;

constexpr uint64_t kMaxProfileRange = 1ull << 27;

std::shared_ptr<Profile> profile_of(const TablePtr& t, int kcol, const std::vector<int>& attrs, hipStream_t s) {
  ProfileKey key(t.get(), kcol, attrs);
  {
    std::lock_guard<std::mutex> g(g_prof_mu);
    auto it = g_profiles.find(key);
    if (it != g_profiles.end() && it->second->owner.lock() == t && it->second->version == t->data_version())
      return it->second;
  }
  auto p = std::make_shared<Profile>();
  p->owner = t;
  p->version = t->data_version();
  p->attrs = attrs;
  KeyEntry ke = table_key(t, kcol, s);
  const uint64_t n = t->rows;
  const uint64_t range = n ? (uint64_t)((int64_t)ke.kmax - (int64_t)ke.kmin + 1) : 0;
  int total = 0;
  for (int a : attrs) {
    auto d = table_dict(t, a, s);
    p->dicts.push_back(d);
    p->shift.push_back(total);
    p->bits.push_back(bits_for(d->n));
    total += p->bits.back();
  }
  p->kmin = ke.kmin;
  p->range = (uint32_t)std::min<uint64_t>(range, kMaxProfileRange);
  if (ke.unique && total <= 31 && range <= kMaxProfileRange && (int)attrs.size() <= bk::kMaxAttr) {
    if (n) {
      bk::RawArgs ra;
      memset(&ra, 0, sizeof(ra));
      ra.nattr = (int)attrs.size();
      for (size_t k = 0; k < attrs.size(); ++k) {
        ra.codes[k] = p->dicts[k]->codes.as<int32_t>();
        ra.shift[k] = p->shift[k];
      }
      DevBuf raw(n * 4 + 4), pcode(n * 4 + 4);
      bk::launch_profile_raw(ra, n, raw.as<int32_t>(), s);
      DevBuf praw(n * 4 + 4);
      p->P = (uint32_t)k::dict_build(raw.as<int32_t>(), n, praw.as<int32_t>(), s);
      if (p->P <= bk::kMaxProfilesPerSlot) {
        p->narrow = p->P < 0xFFu;
        const uint64_t pbytes = p->narrow ? range : range * 2;
        p->pkey = DevBuf(pbytes + 256);
        // keys without a dimension row (and the entry past the range) hold P
        const uint32_t fill = p->narrow ? p->P * 0x01010101u : p->P | p->P << 16;
        k::fill_u32(p->pkey.p, pbytes / 4 + 2, fill, s);
        k::codes_of(raw.as<int32_t>(), n, praw.as<int32_t>(), p->P, pcode.as<int32_t>(), s);
        bk::launch_profile_scatter((const int32_t*)t->col_ptr(kcol), n, ke.kmin, pcode.as<int32_t>(), p->pkey.p,
                                   p->narrow ? 1 : 0, s);
        p->praw = DevBuf((uint64_t)p->P * 4 + 4);
        SVM_HIP(hipMemcpyAsync(p->praw.p, praw.p, (uint64_t)p->P * 4, hipMemcpyDeviceToDevice, s));
        p->ok = true;
      }
      SVM_HIP(hipStreamSynchronize(s));
    } else {
      p->pkey = DevBuf(256);
      k::fill_u32(p->pkey.p, 64, 0u, s);  // P = 0
      p->praw = DevBuf(4);
      p->ok = true;
    }
  }
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_profiles[key] = p;
  return p;
}

// ------------------------------------------------------------------ lowering
struct JoinL {
  TablePtr dim;
  int kcol = 0, fcol = 0;
  std::vector<Atom> atoms;  // on dimension columns
  int key = -1;             // key component read through this join
  int attr = -1;            // its dimension column
  std::vector<int> attrs;   // atom columns and the key column, ascending
};
struct KeyL {
  int join = 0;
  std::shared_ptr<DictEntry> dict;
  uint32_t stride = 1;
};
struct FactLowered {
  bool ok = false;
  TablePtr src, base;
  long long row0 = 0, row1 = 0;
  std::vector<std::array<int, 3>> preds;  // fact column, lo, hi
  int mop = 0, ma = 0, mb = 0;
  int K = 0;
  uint64_t G = 1;
  std::vector<KeyL> keys;
  std::vector<JoinL> joins;
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
  if (!info.isGPU || !p.is_reduce || !p.select_exp.empty() || p.has_loader || info.pre_pipeline_id.empty()) return f;
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
  f.row0 = (long long)src->row0;
  f.row1 = (long long)(src->row0 + src->rows);
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
    if (kv.second.first > kv.second.second) f.preds.push_back({kv.first, 1, 0});
    else
      f.preds.push_back({kv.first, (int)std::max<long long>(kv.second.first, INT_MIN),
                         (int)std::min<long long>(kv.second.second, INT_MAX)});
  }
  // metric
  if (!parse_metric(p.metric_exp[0], f.mop, f.ma, f.mb) || !src_int(f.ma) || !src_int(f.mb)) return f;
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
  f.keys.resize(K);
  for (int k = (int)K - 1; k >= 0; --k) {
    f.keys[k].dict = table_dict(refs[key_obj[k] - 1].base, key_base[k], s);
    f.keys[k].stride = (uint32_t)G;
    G *= f.keys[k].dict->n;
    if (G > (1ull << 22)) return f;
  }
  f.K = (int)K;
  f.G = K ? G : 1;
  // joins
  std::set<int> seen;
  std::set<std::tuple<const Table*, int, int>> slots;
  uint64_t profiles = 0;
  for (size_t j = 0; j < p.join_rkey.size(); ++j) {
    KeyIndex L = p.join_lkey[j], R = p.join_rkey[j];
    if (L.obj != 0 || !src_int(L.col) || R.obj < 1 || R.obj > (int)refs.size()) return f;
    if (!seen.insert(R.obj).second) return f;
    const DimLowered& d = refs[R.obj - 1];
    if (R.col < 0 || R.col >= (int)d.proj.size()) return f;
    JoinL jl;
    jl.dim = d.base;
    jl.kcol = d.proj[R.col];
    jl.fcol = L.col;
    if (d.base->cols[jl.kcol].type != Int32) return f;
    // the absorbed dimension object is the filtered base; it is all in bulk 0 only
    // when the base is (a join reads bulk 0 alone, core/pipeline.cu:91)
    if (d.base->bulk0_rows() < d.base->rows) return f;
    // When the base's own hash table would drop inserts, a query's dimension object (an
    // in-order subset of the base's rows) may drop some too: the pass checks each query
    // and leaves those that drop to the plain path.  Clusters wider than the check
    // replays keep the query on the plain path from the start.
    if (auto hc = table_drop_clusters(d.base, jl.kcol, pipeline_ht_slots(E), s))
      if (hc->max_span > (uint32_t)bk::kCheckSpan) return f;
    if (!slots.insert(std::make_tuple(d.base.get(), jl.kcol, jl.fcol)).second) return f;
    jl.atoms = d.atoms;
    std::set<int> attrs;
    for (auto& a : jl.atoms) {
      attrs.insert(a.col);
      if (a.kind != bk::DP_RANGE) {
        auto dict = table_dict(d.base, a.col, s);
        if (!string_rep(d.base, a.col, dict, s)->exact) return f;
      }
    }
    for (size_t k = 0; k < K; ++k)
      if (key_obj[k] == R.obj) {
        jl.key = (int)k;
        jl.attr = key_base[k];
        attrs.insert(jl.attr);
        f.keys[k].join = (int)j;
      }
    jl.attrs.assign(attrs.begin(), attrs.end());
    auto prof = profile_of(d.base, jl.kcol, jl.attrs, s);
    if (!prof->ok) return f;
    profiles += prof->P + 1;  // and the no-row profile
    f.joins.push_back(std::move(jl));
  }
  if (profiles > (uint64_t)bk::kMaxProfiles) return f;
  // every key component's dimension must be joined
  for (size_t k = 0; k < K; ++k)
    if (!seen.count(key_obj[k])) return f;
  f.src = src;
  f.base = base;
  f.ok = true;
  return f;
}

}  // namespace

// Fixed layout of one pass's device descriptors: every pointer a launch of the
// pass receives stays the same from batch to batch, so a captured launch
// sequence replays for any batch once its descriptors are uploaded in place.
constexpr size_t al(size_t b) { return (b + 255) & ~size_t(255); }
struct PassLayout {
  static constexpr size_t o_args = 0;
  static constexpr size_t o_margs = al(sizeof(bk::ScanArgs));
  static constexpr size_t o_qd = o_margs + al(sizeof(bk::MaskArgs));
  static constexpr size_t o_row = o_qd + al(bk::kMaxBatch * sizeof(bk::QDesc));
  static constexpr size_t o_terms = o_row + al(bk::kMaxBatch * 16);
  static constexpr size_t o_emit = o_terms + al(bk::kMaxTerms * sizeof(bk::MaskTerm));
  static constexpr size_t o_chk = o_emit + al(bk::kMaxBatch * sizeof(bk::Emit));
  static constexpr size_t total = o_chk + al(bk::kMaxSlot * sizeof(bk::DropCheck));
};
// Emission blocks of one pass: 256 flag words per block, G <= 2^22 per query.
constexpr uint64_t kMaxEmitBlocks = (uint64_t)bk::kMaxBatch * (((1u << 22) / 32 + 255) / 256);

// Device buffers and captured launch sequences reused across batches (buffers are
// grown on demand and never shrunk), so a steady stream of batches allocates
// nothing and launches one graph per pass.
struct Scratch {
  std::vector<DevBuf> bufs;
  void* host = nullptr;  // pinned staging for descriptors
  size_t host_bytes = 0;
  // (variant, LDS profiles, grid, drop check, key bits)
  std::map<std::tuple<int, unsigned, int, bool, bool>, hipGraphExec_t> graphs;
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
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (host) (void)hipHostFree(host);
  }
};

// HBM a fused query holds while it runs: its group table and its share of the
// batch's profile masks (profiles themselves are cached with the dimension tables).
static uint64_t fused_bytes(const FactLowered& f) {
  return 4096 + f.G * 8 + (f.G + 31) / 32 * 4 + (uint64_t)bk::kMaxProfiles * 8 / bk::kMaxBatch;
}

namespace {

// One shared scan: at most kMaxBatch queries whose joins fit kMaxSlot slots.
struct SlotB {
  TablePtr dim;
  int kcol = 0, fcol = 0;
  std::set<int> attrs;
  std::shared_ptr<Profile> prof;
};
struct PassB {
  std::vector<int> qi;
  std::vector<SlotB> slots;
  std::vector<std::array<int, 3>> metrics;
  int npred = 0;
};

int slot_bits(const SlotB& sl, hipStream_t s) {
  int b = 0;
  for (int a : sl.attrs) b += bits_for(table_dict(sl.dim, a, s)->n);
  return b;
}

// Adds query i to the pass if every limit of one shared scan still holds.
bool pass_add(PassB& P, const FactLowered& f, int i, hipStream_t s) {
  if ((int)P.qi.size() >= bk::kMaxBatch) return false;
  if (P.npred + (int)f.preds.size() > bk::kMaxFactPred) return false;
  PassB T = P;
  for (auto& j : f.joins) {
    SlotB* sl = nullptr;
    for (auto& x : T.slots)
      if (x.dim == j.dim && x.kcol == j.kcol && x.fcol == j.fcol) sl = &x;
    if (!sl) {
      if ((int)T.slots.size() >= bk::kMaxSlot) return false;
      T.slots.push_back(SlotB{j.dim, j.kcol, j.fcol, {}, nullptr});
      sl = &T.slots.back();
    }
    sl->attrs.insert(j.attrs.begin(), j.attrs.end());
    if ((int)sl->attrs.size() > bk::kMaxAttr || slot_bits(*sl, s) > 31) return false;
  }
  std::array<int, 3> m = {f.mop, f.ma, f.mop == bk::MOP_COL ? f.ma : f.mb};
  if (std::find(T.metrics.begin(), T.metrics.end(), m) == T.metrics.end()) {
    if ((int)T.metrics.size() >= bk::kMaxMetric) return false;
    T.metrics.push_back(m);
  }
  T.npred += (int)f.preds.size();
  T.qi.push_back(i);
  P = std::move(T);
  return true;
}

// Rebuilds a pass from a subset of its queries.
PassB pass_of(const std::vector<FactLowered*>& qs, const std::vector<int>& idx, hipStream_t s) {
  PassB P;
  for (int i : idx)
    if (!pass_add(P, *qs[i], i, s)) fail(SVM_ERR_INVALID, "shared scan limits violated by a subset");
  return P;
}

}  // namespace

namespace {

// Autotuner of the shared scan: per batch shape (slots, queries, rows scanned) the
// fastest (variant, tile rows) among the candidates, timed on the batch itself the
// first time the shape is seen.  A trial runs the whole scan; the pass is cleared
// again before the run whose results are kept.
struct BatchTune {
  int variant = 0;
  long long tile = 16384;
  int bits = 0;
  // trials of the tuning run: (variant, tile, bits, ms); ms < 0 marks a trial whose
  // group tables differed from the first trial's (never chosen)
  std::vector<std::array<double, 4>> trials;
};
std::mutex g_btune_mu;
std::map<std::tuple<int, int, int, bool>, BatchTune> g_btune;
constexpr long long kTileChoices[] = {16384, 65536};
constexpr uint64_t kBatchTuneMinRows = 1ull << 22;

// SVM_BATCH_TUNE_MIN_ROWS=<n>: tune batch shapes from n scanned rows on.
uint64_t batch_tune_min_rows() {
  const char* e = getenv("SVM_BATCH_TUNE_MIN_ROWS");
  if (e && *e) {
    char* end = nullptr;
    const unsigned long long v = strtoull(e, &end, 10);
    if (end && *end == '\0') return v;
  }
  return kBatchTuneMinRows;
}

// SVM_BATCH_VARIANT=<variant>,<tile rows>,<key bits>: every pass runs this shared-scan
// variant (0..scan_variants()-1), tile (16384 or 65536) and key-bit pre-test (0 or 1),
// with no trials.  A variant with 32-bit query masks is used only by passes of at
// most 32 queries (larger passes keep the default).  Returns false when unset.
bool forced_batch_variant(BatchTune& t) {
  const char* e = getenv("SVM_BATCH_VARIANT");
  if (!e || !*e) return false;
  int v = -1, bits = -1;
  long long tile = 0;
  char extra = 0;
  if (sscanf(e, "%d,%lld,%d%c", &v, &tile, &bits, &extra) != 3 || v < 0 || v >= bk::scan_variants() ||
      (tile != kTileChoices[0] && tile != kTileChoices[1]) || (bits != 0 && bits != 1))
    fail(SVM_ERR_INVALID, "SVM_BATCH_VARIANT must be <variant 0.." + std::to_string(bk::scan_variants() - 1) +
                              ">,<16384|65536>,<0|1>");
  t.variant = v;
  t.tile = tile;
  t.bits = bits;
  return true;
}

int ilog2(uint64_t v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    r++;
  }
  return r;
}

unsigned prof_cap(unsigned nprof) { return nprof <= 1024 ? 1024u : nprof <= 2048 ? 2048u : 4096u; }

// Persistent grid: as many workgroups per CU as LDS and 20 waves per CU (five per
// SIMD at the scan's register count) allow.
int scan_grid(const Engine& E, int variant, unsigned cap) {
  const size_t lds = bk::shared_scan_lds(variant, cap) + 64;
  const int by_lds = (int)((160 * 1024) / lds);
  const int by_waves = 20 / (bk::scan_variant_block(variant) / 64);
  return E.num_cus * std::max(1, std::min(by_lds, std::min(by_waves, 8)));
}

bool env_off(const char* name) {
  const char* v = getenv(name);
  return v && v[0] == '0';
}

}  // namespace

std::string batch_tune_report(bool reset) {
  std::lock_guard<std::mutex> g(g_btune_mu);
  std::string j = "[";
  bool first = true;
  for (auto& kv : g_btune) {
    const BatchTune& t = kv.second;
    int rejected = 0;
    std::string tr = "[";
    for (size_t q = 0; q < t.trials.size(); ++q) {
      rejected += t.trials[q][3] < 0 ? 1 : 0;
      char b[96];
      snprintf(b, sizeof(b), "%s[%d,%lld,%d,%.4f]", q ? "," : "", (int)t.trials[q][0], (long long)t.trials[q][1],
               (int)t.trials[q][2], t.trials[q][3]);
      tr += b;
    }
    char h[160];
    snprintf(h, sizeof(h), "{\"slots\":%d,\"log2_queries\":%d,\"log2_rows\":%d,\"mask32\":%s,\"variant\":%d,"
                           "\"tile\":%lld,\"bits\":%d,", std::get<0>(kv.first), std::get<1>(kv.first),
             std::get<2>(kv.first), std::get<3>(kv.first) ? "true" : "false", t.variant, t.tile, t.bits);
    j += std::string(first ? "" : ",") + h + "\"rejected\":" + std::to_string(rejected) + ",\"trials\":" + tr + "]}";
    first = false;
  }
  if (reset) g_btune.clear();
  return j + "]";
}

static void run_pass(const std::vector<FactLowered*>& qs, PassB& P,
                     std::vector<std::unique_ptr<svm_result>*>& outs, Engine& E, hipStream_t s, Scratch& sc,
                     bool tune, std::vector<int>& redo) {
  bk::prepare_shared_scan();
  TablePtr base = qs[P.qi[0]]->base;
  const int nq = (int)P.qi.size();
  // slots, with the profile of their attribute union
  unsigned nprof = 0;
  for (auto& sl : P.slots) {
    std::vector<int> attrs(sl.attrs.begin(), sl.attrs.end());
    sl.prof = profile_of(sl.dim, sl.kcol, attrs, s);
    if (!sl.prof->ok) fail(SVM_ERR_UNSUPPORTED, "dimension profile unavailable");
    nprof += sl.prof->P + 1;
  }
  // Slots are looked up in order, each gated by the ones before it (a row no query
  // still wants reads the shared no-row entry), so the order sets how many rows
  // reach each lookup.  As for any chain of filters, the slots go in ascending order
  // of cost / (1 - pass fraction): the cost of a lookup by the size of its
  // key-to-profile array (up to 64 KB it stays in LDS or L2 at little cost, up to
  // 2 MB in L2; larger ones miss L2 for most keys), the pass fraction estimated as
  // that of keys passing some joining query's filter (an equality atom passes 1/n of
  // a column's n distinct values, an inequality 1 - 1/n, a range about half; queries
  // taken as independent).  The masks are ANDed, so the order does not change any
  // result.
  // estimated fraction of a slot's keys that pass join j's dimension filter
  auto join_pass = [&](const JoinL& j) {
    double f = 1.0;
    for (const Atom& a : j.atoms) {
      const double n = std::max<uint32_t>(1u, table_dict(j.dim, a.col, s)->n);
      f *= a.kind == bk::DP_STREQ ? 1.0 / n : a.kind == bk::DP_STRNE ? 1.0 - 1.0 / n
           : a.lo > a.hi          ? 0.0
           : a.lo == a.hi         ? 1.0 / n
                                  : 0.5;
    }
    return f;
  };
  {
    std::vector<std::pair<double, uint64_t>> key(P.slots.size());
    for (size_t k = 0; k < P.slots.size(); ++k) {
      const SlotB& sl = P.slots[k];
      const uint64_t entries = (uint64_t)sl.prof->range + 1;
      const unsigned ebits = sl.prof->P + 1 <= 16 ? 4 : sl.prof->P + 1 <= 256 ? 8 : 16;
      double miss = 1.0;  // no joining query passes
      for (int i : P.qi)
        for (const JoinL& j : qs[i]->joins) {
          if (j.dim != sl.dim || j.kcol != sl.kcol || j.fcol != sl.fcol) continue;
          miss *= 1.0 - join_pass(j);
        }
      const uint64_t bytes = entries * ebits / 8;
      const double cost = bytes <= (64u << 10) ? 0.25 : bytes <= (2u << 20) ? 1.0 : 3.0;
      key[k] = {cost / std::max(1e-3, miss), bytes};
    }
    std::vector<size_t> ord(P.slots.size());
    for (size_t k = 0; k < ord.size(); ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
    std::vector<SlotB> sorted;
    for (size_t k : ord) sorted.push_back(std::move(P.slots[k]));
    P.slots = std::move(sorted);
  }
  if (nprof > (unsigned)bk::kMaxProfiles) fail(SVM_ERR_INVALID, "pass exceeds the LDS profile budget");
  auto slot_index = [&](const JoinL& j) {
    for (size_t k = 0; k < P.slots.size(); ++k)
      if (P.slots[k].dim == j.dim && P.slots[k].kcol == j.kcol && P.slots[k].fcol == j.fcol) return (int)k;
    fail(SVM_ERR_INVALID, "join without a slot");
    return -1;
  };
  auto attr_pos = [&](const Profile& pr, int col) {
    for (size_t k = 0; k < pr.attrs.size(); ++k)
      if (pr.attrs[k] == col) return (int)k;
    fail(SVM_ERR_INVALID, "attribute missing from a profile");
    return -1;
  };
  // buffers: group tables grow with the batch, everything else has a fixed size
  uint64_t sum_words = 0, flag_words = 0;
  for (int i : P.qi) {
    sum_words += qs[i]->G;
    flag_words += (qs[i]->G + 31) / 32;
  }
  unsigned long long* masks = (unsigned long long*)sc.get(0, (uint64_t)bk::kMaxProfiles * 8);
  unsigned long long* sums = (unsigned long long*)sc.get(1, sum_words * 8 + 8);
  unsigned* flags = (unsigned*)sc.get(2, flag_words * 4 + 4);
  unsigned* counter = (unsigned*)sc.get(3, 16);
  char* dd = (char*)sc.get(4, PassLayout::total);
  unsigned long long* counts = (unsigned long long*)sc.get(5, kMaxEmitBlocks * 8);
  unsigned long long* offsets = (unsigned long long*)sc.get(6, kMaxEmitBlocks * 8);
  unsigned long long* totals = (unsigned long long*)sc.get(7, bk::kMaxBatch * 8);
  uint64_t bit_words = 0;
  for (auto& sl : P.slots) bit_words += ((uint64_t)sl.prof->range + 63) / 64 * 2;
  unsigned* kbits = (unsigned*)sc.get(10, bit_words * 4 + 8);
  const auto* d_args = (const bk::ScanArgs*)(dd + PassLayout::o_args);
  const auto* d_margs = (const bk::MaskArgs*)(dd + PassLayout::o_margs);
  const auto* d_emit = (const bk::Emit*)(dd + PassLayout::o_emit);
  // descriptors, written in place in the pinned staging area
  char* host = sc.pinned(PassLayout::total);
  auto* A = (bk::ScanArgs*)(host + PassLayout::o_args);
  auto* M = (bk::MaskArgs*)(host + PassLayout::o_margs);
  auto* qd = (bk::QDesc*)(host + PassLayout::o_qd);
  auto* qrow = (long long*)(host + PassLayout::o_row);
  auto* terms = (bk::MaskTerm*)(host + PassLayout::o_terms);
  auto* em = (bk::Emit*)(host + PassLayout::o_emit);
  memset(host, 0, PassLayout::total);
  A->nq = nq;
  A->nslot = (int)P.slots.size();
  A->nmetric = (int)P.metrics.size();
  for (size_t c = 0; c < base->cols.size() && c < (size_t)bk::kMaxCols; ++c) A->cols[c] = base->cols[c].dev;
  unsigned lds = 0;
  for (size_t k = 0; k < P.slots.size(); ++k) {
    const SlotB& sl = P.slots[k];
    bk::Slot& S = A->slot[k];
    S.fcol = sl.fcol;
    S.kmin = sl.prof->kmin;
    S.range = sl.prof->range;
    S.P = sl.prof->P;
    S.pkey = sl.prof->pkey.p;
    S.narrow = sl.prof->narrow ? 1 : 0;
    S.praw = sl.prof->praw.as<int32_t>();
    S.mask = masks + lds;
    S.bits = kbits;
    kbits += ((uint64_t)S.range + 63) / 64 * 2;
    S.lds = (int)lds;
    lds += sl.prof->P + 1;
  }
  for (size_t m = 0; m < P.metrics.size(); ++m) {
    A->mop[m] = P.metrics[m][0];
    A->ma[m] = P.metrics[m][1];
    A->mb[m] = P.metrics[m][2];
  }
  std::vector<std::vector<bk::FactPred>> by_col(bk::kMaxCols);
  std::vector<std::vector<bk::MaskTerm>> slot_terms(P.slots.size());
  long long row_lo = LLONG_MAX, row_hi = 0;
  uint64_t so = 0, fo = 0, blocks = 0;
  for (int q = 0; q < nq; ++q) {
    const FactLowered& f = *qs[P.qi[q]];
    qrow[2 * q] = f.row0;
    qrow[2 * q + 1] = f.row1;
    if (f.row1 > f.row0) {
      row_lo = std::min(row_lo, f.row0);
      row_hi = std::max(row_hi, f.row1);
    }
    unsigned qcols = 1u << f.ma | 1u << f.mb;
    for (auto& pr : f.preds) {
      by_col[pr[0]].push_back(bk::FactPred{q, pr[1], pr[2]});
      qcols |= 1u << pr[0];
    }
    for (auto& j : f.joins) qcols |= 1u << j.fcol;
    A->qcols[q] = qcols;
    bk::QDesc& d = qd[q];
    d.K = f.K;
    std::array<int, 3> m = {f.mop, f.ma, f.mop == bk::MOP_COL ? f.ma : f.mb};
    d.metric = (int)(std::find(P.metrics.begin(), P.metrics.end(), m) - P.metrics.begin());
    d.sum_off = (unsigned)so;
    d.flag_off = (unsigned)fo;
    d.G = (unsigned)f.G;
    d.dense = bk::kNoDense;
    for (auto& j : f.joins) {
      const int sk = slot_index(j);
      A->slot[sk].joined |= 1ull << q;
      const Profile& pr = *P.slots[sk].prof;
      bk::MaskTerm t;
      memset(&t, 0, sizeof(t));
      t.q = q;
      for (auto& a : j.atoms) {
        bk::MaskAtom& ma = t.atom[t.natom++];
        const int pos = attr_pos(pr, a.col);
        ma.kind = a.kind;
        ma.shift = pr.shift[pos];
        ma.bits = pr.bits[pos];
        ma.lo = (int)std::max<long long>(a.lo, INT_MIN);
        ma.hi = (int)std::min<long long>(a.hi, INT_MAX);
        if (a.lo > a.hi) { ma.lo = 1; ma.hi = 0; }
        ma.n = a.n;
        ma.dict = pr.dicts[pos]->dict.as<int32_t>();
        if (a.kind != bk::DP_RANGE) {
          ma.col = (const char*)j.dim->col_ptr(a.col);
          ma.size = (int)j.dim->cols[a.col].size;
          ma.rep = string_rep(j.dim, a.col, pr.dicts[pos], s)->rep.as<int32_t>();
        }
        memcpy(ma.lit, a.lit.data(), std::min<size_t>(a.lit.size(), sizeof(ma.lit) - 1));
      }
      slot_terms[sk].push_back(t);
      if (j.key >= 0) {
        const int pos = attr_pos(pr, j.attr);
        d.key[j.key] = (unsigned)A->slot[sk].lds | (unsigned)sk << 13 | (unsigned)pr.shift[pos] << 16 |
                       (unsigned)pr.bits[pos] << 21;
        d.stride[j.key] = f.keys[j.key].stride;
      }
    }
    bk::Emit& e = em[q];
    e.flags = flags + fo;
    e.sums = sums + so;
    e.G = (unsigned)f.G;
    e.K = f.K;
    for (int k = 0; k < f.K; ++k) {
      e.n[k] = f.keys[k].dict->n;
      e.stride[k] = f.keys[k].stride;
      e.dict[k] = f.keys[k].dict->dict.as<int32_t>();
    }
    e.block0 = blocks;
    e.nblocks = ((f.G + 31) / 32 + 255) / 256;
    blocks += e.nblocks;
    so += f.G;
    fo += (f.G + 31) / 32;
  }
  int np = 0;
  for (int c = 0; c < bk::kMaxCols; ++c) {
    A->pred0[c] = np;
    for (auto& fp : by_col[c]) A->pred[np++] = fp;
  }
  A->pred0[bk::kMaxCols] = np;
  // value segments of each column's predicates (one search per row instead of one
  // test per predicate when a tile has many active predicates on the column)
  {
    int nb = 0;
    for (int c = 0; c < bk::kMaxCols; ++c) {
      A->seg0[c] = nb;
      std::vector<long long> bounds;
      unsigned long long cq = 0;
      for (auto& fp : by_col[c]) {
        cq |= 1ull << fp.q;
        if (fp.lo > INT_MIN) bounds.push_back(fp.lo);
        if (fp.hi < INT_MAX) bounds.push_back((long long)fp.hi + 1);
      }
      std::sort(bounds.begin(), bounds.end());
      bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
      for (size_t j = 0; j <= bounds.size(); ++j) {
        // every value of segment j (from its lower bound on) meets the same predicates
        const long long v = j == 0 ? (long long)INT_MIN : bounds[j - 1];
        unsigned long long m = ~0ull;
        for (auto& fp : by_col[c])
          if (v < fp.lo || v > fp.hi) m &= ~(1ull << fp.q);
        A->segm[nb + c + (int)j] = m;
      }
      for (long long b : bounds) A->segb[nb++] = (int)b;
      A->colq[c] = cq;
      int steps = 0;
      while ((1u << steps) <= bounds.size()) ++steps;
      A->segsteps[c] = steps;
    }
    A->seg0[bk::kMaxCols] = nb;
  }
  // Staged variants: the key columns of the first slots form tier 0 when that is
  // estimated to read fewer bytes.  After slots 0..k-1 a row passes query q with the
  // product of q's estimated join pass fractions on them (times the share of the
  // union of slices q covers), and passes some query with 1 - prod_q (1 - that);
  // every other column is then read in the 128-byte lines (32 rows) holding such a
  // row.  k = 0 (every column for every row) unless some k costs at most 85 % of it.
  {
    unsigned cols = 0;
    for (int q = 0; q < nq; ++q) cols |= A->qcols[q];
    const double span = row_hi > row_lo ? (double)(row_hi - row_lo) : 1.0;
    const double all = (double)__builtin_popcount(cols);
    double best = 0.85 * all;
    A->nslot0 = 0;
    A->tier0 = 0x3Fu;
    for (size_t k = 1; k <= P.slots.size(); ++k) {
      unsigned t0 = 0;
      for (size_t sk = 0; sk < k; ++sk) t0 |= 1u << P.slots[sk].fcol;
      double none = 1.0;  // a row passes no query
      for (int q = 0; q < nq; ++q) {
        const FactLowered& f = *qs[P.qi[q]];
        double pq = (double)(f.row1 - f.row0) / span;
        for (auto& j : f.joins)
          if (slot_index(j) < (int)k) pq *= join_pass(j);
        none *= 1.0 - std::min(1.0, std::max(0.0, pq));
      }
      const double line = 1.0 - std::pow(none, 32.0);
      const double cost = (double)__builtin_popcount(cols & t0) + (double)__builtin_popcount(cols & ~t0) * line;
      if (cost < best) {
        best = cost;
        A->nslot0 = (int)k;
        A->tier0 = t0;
      }
    }
    // SVM_BATCH_TIER0_SLOTS=<k>: tier 0 holds the first k slots' key columns whatever
    // the estimate (0: no staging)
    if (const char* e = getenv("SVM_BATCH_TIER0_SLOTS")) {
      char* end = nullptr;
      const long k = strtol(e, &end, 10);
      if (!*e || !end || *end || k < 0) fail(SVM_ERR_INVALID, "SVM_BATCH_TIER0_SLOTS must be a slot count");
      A->nslot0 = (int)std::min<long>(k, (long)P.slots.size());
      A->tier0 = 0x3Fu;
      if (A->nslot0 > 0) {
        A->tier0 = 0;
        for (int sk = 0; sk < A->nslot0; ++sk) A->tier0 |= 1u << P.slots[sk].fcol;
      }
    }
  }
  int nt = 0;
  for (size_t k = 0; k < P.slots.size(); ++k) {
    bk::MaskSlot& ms = M->slots[k];
    ms.P = A->slot[k].P;
    ms.praw = A->slot[k].praw;
    ms.mask = masks + A->slot[k].lds;
    ms.t0 = nt;
    for (auto& t : slot_terms[k]) terms[nt++] = t;
    ms.t1 = nt;
  }
  M->terms = (const bk::MaskTerm*)(dd + PassLayout::o_terms);
  M->nslot = (int)P.slots.size();
  // slots whose dimension table drops inserts in the reference's hash table
  auto* chk = (bk::DropCheck*)(host + PassLayout::o_chk);
  std::vector<std::shared_ptr<const k::HtClusters>> held;
  int nchk = 0;
  for (size_t k = 0; k < P.slots.size(); ++k) {
    auto hc = table_drop_clusters(P.slots[k].dim, P.slots[k].kcol, pipeline_ht_slots(E), s);
    if (!hc) continue;
    if (hc->max_span > (uint32_t)bk::kCheckSpan) fail(SVM_ERR_INVALID, "drop check wider than its replay window");
    held.push_back(hc);
    const bk::Slot& S = A->slot[k];
    bk::DropCheck& D = chk[nchk++];
    D.start = hc->start.as<unsigned>();
    D.key = hc->key.as<int>();
    D.home = hc->home.as<unsigned>();
    D.clusters = hc->clusters;
    D.kmin = S.kmin;
    D.range = S.range;
    D.P = S.P;
    D.pkey = S.pkey;
    D.narrow = S.narrow;
    D.mask = S.mask;
    D.joined = S.joined;
  }
  unsigned long long* dropq = (unsigned long long*)sc.get(11, 16);
  A->chk = nchk ? (const bk::DropCheck*)(dd + PassLayout::o_chk) : nullptr;
  A->nchk = nchk;
  A->dropq = dropq;
  A->nprof = nprof;
  A->qd = (const bk::QDesc*)(dd + PassLayout::o_qd);
  A->qrow = (const long long*)(dd + PassLayout::o_row);
  A->sums = sums;
  A->flags = flags;
  A->sum_words = sum_words;
  A->flag_words = flag_words;
  A->total_blocks = blocks;
  A->counter = counter;
  if (row_lo == LLONG_MAX) row_lo = row_hi = 0;
  A->row_lo = row_lo & ~3ll;
  A->row_hi = row_hi;
  const unsigned cap = prof_cap(nprof);
  auto set_tile = [&](long long tile) {
    A->tile_rows = tile;
    A->ntiles = (unsigned)((row_hi - A->row_lo + tile - 1) / tile);
  };
  // autotuner: the launch shape of this batch shape
  const uint64_t union_rows = (uint64_t)(row_hi - A->row_lo);
  const bool mask32 = nq <= 32;  // query masks of 32 bits suffice
  const auto tkey = std::make_tuple(A->nslot, ilog2((uint64_t)nq), ilog2(union_rows + 1), mask32);
  BatchTune ch;
  // untuned: 512 threads, LDS hash, gated lookups; at most 32 queries: 1,024 threads,
  // dense tables, LDS maps and the row queue (selective passes such as whole-table
  // star joins spend least time there)
  ch.variant = mask32 ? bk::kDefaultVariantMask32 : 0;
  bool tuned = false;
  {
    std::lock_guard<std::mutex> g(g_btune_mu);
    auto it = g_btune.find(tkey);
    if (it != g_btune.end()) {
      ch = it->second;
      tuned = true;
    }
  }
  if (!tuned) ch.tile = union_rows > (1ull << 26) ? 65536 : 16384;
  BatchTune forced;
  if (forced_batch_variant(forced) && (!bk::scan_variant_mask32(forced.variant) || mask32)) {
    ch = forced;
    tuned = true;
  }
  if (!tuned && tune && union_rows >= batch_tune_min_rows() && !env_off("SVM_AUTOTUNE")) {
    hipEvent_t t0, t1;
    SVM_HIP(hipEventCreate(&t0));
    SVM_HIP(hipEventCreate(&t1));
    // Integer results do not depend on the launch shape: the first trial's group
    // tables are kept and every later trial's are compared with them on the device;
    // a trial that differs is never chosen.
    uint64_t* ref_sums = (uint64_t*)sc.get(12, sum_words * 8 + 8);
    unsigned* ref_flags = (unsigned*)sc.get(13, flag_words * 4 + 4);
    uint32_t* diff = (uint32_t*)sc.get(14, 16);
    std::vector<std::array<double, 4>> trials;
    auto trial = [&](const BatchTune& c) {
      set_tile(c.tile);
      A->use_bits = c.bits;
      if (bk::scan_variant_dense(c.variant)) bk::assign_dense(qd, nq, bk::scan_variant_hash_slots(c.variant));
      bk::assign_lds_maps(A, c.variant, cap);
      SVM_HIP(hipMemcpyAsync(dd, host, PassLayout::total, hipMemcpyHostToDevice, s));
      bk::launch_pass_clear(d_args, s);
      bk::launch_profile_masks(d_margs, s);
      bk::launch_slot_bits(d_args, s);
      SVM_HIP(hipEventRecord(t0, s));
      bk::launch_shared_scan(d_args, c.variant, cap, scan_grid(E, c.variant, cap), s);
      SVM_HIP(hipEventRecord(t1, s));
      SVM_HIP(hipStreamSynchronize(s));
      float ms = 0;
      SVM_HIP(hipEventElapsedTime(&ms, t0, t1));
      if (trials.empty()) {
        SVM_HIP(hipMemcpyAsync(ref_sums, sums, sum_words * 8, hipMemcpyDeviceToDevice, s));
        SVM_HIP(hipMemcpyAsync(ref_flags, flags, flag_words * 4, hipMemcpyDeviceToDevice, s));
      } else if (k::words_differ(ref_sums, sums, sum_words * 2, diff, s) ||
                 k::words_differ(ref_flags, flags, flag_words, diff, s)) {
        fprintf(stderr, "[svm] batch tune: variant %d, tile %lld, key bits %d gave different group tables; not "
                        "chosen\n", c.variant, c.tile, c.bits);
        ms = -1.0f;
      }
      trials.push_back({(double)c.variant, (double)c.tile, (double)c.bits, (double)ms});
      if (E.cfg.verbose)
        fprintf(stderr, "[svm] batch tune: %d threads, %d LDS slots (%s), %s lookups%s%s, key bits %d, tile %lld rows: %.3f ms\n",
                bk::scan_variant_block(c.variant), bk::scan_variant_hash_slots(c.variant),
                bk::scan_variant_dense(c.variant) ? "dense" : "hash",
                bk::scan_variant_parallel(c.variant) ? "parallel" : "gated",
                bk::scan_variant_prefetch(c.variant) ? ", prefetch" : "",
                bk::scan_variant_staged(c.variant) ? (", staged over " + std::to_string(A->nslot0) + " slots").c_str() : "",
                c.bits, c.tile, ms);
      return ms < 0 ? 1e30f : ms;
    };
    float best = 1e30f;
    ch.variant = -1;
    for (int v = 0; v < bk::scan_variants(); ++v)
      for (long long tile : kTileChoices) {
        if (bk::scan_variant_mask32(v) != mask32) continue;
        BatchTune c{v, tile, 0};
        const float ms = trial(c);
        if (ms < best) {
          best = ms;
          ch = c;
        }
      }
    if (ch.variant < 0) fail(SVM_ERR_DEVICE, "batch autotuner: no trial reproduced the first trial's group tables");
    BatchTune c = ch;
    c.bits = 1;
    if (trial(c) < best) ch = c;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    ch.trials = std::move(trials);
    std::lock_guard<std::mutex> g(g_btune_mu);
    g_btune[tkey] = ch;
  }
  A->use_bits = ch.bits;
  set_tile(ch.tile);
  if (bk::scan_variant_dense(ch.variant)) bk::assign_dense(qd, nq, bk::scan_variant_hash_slots(ch.variant));
  bk::assign_lds_maps(A, ch.variant, cap);
  const int grid = scan_grid(E, ch.variant, cap);
  // the key bits are built only when the scan reads them
  bool need_bits = A->use_bits != 0;
  for (int k = 0; k < A->nslot; ++k) need_bits = need_bits || A->slot[k].lbits >= 0;
  // the pass: upload descriptors, then clear -> masks -> scan -> emission count
  hipEvent_t e0, e1, eb, es;
  SVM_HIP(hipEventCreate(&e0));
  SVM_HIP(hipEventCreate(&e1));
  SVM_HIP(hipEventCreate(&eb));
  SVM_HIP(hipEventCreate(&es));
  SVM_HIP(hipMemcpyAsync(dd, host, PassLayout::total, hipMemcpyHostToDevice, s));
  SVM_HIP(hipEventRecord(e0, s));
  const bool graph = !env_off("SVM_BATCH_GRAPH");
  if (graph) {
    hipGraphExec_t& ge = sc.graphs[std::make_tuple(ch.variant, cap, grid, nchk > 0, need_bits)];
    if (!ge) {
      hipGraph_t g = nullptr;
      SVM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      bk::launch_pass_clear(d_args, s);
      bk::launch_profile_masks(d_margs, s);
      if (nchk) bk::launch_drop_check(d_args, s);
      if (need_bits) bk::launch_slot_bits(d_args, s);
      bk::launch_shared_scan(d_args, ch.variant, cap, grid, s);
      bk::launch_emit_count(d_emit, d_args, counts, offsets, totals, s);
      SVM_HIP(hipStreamEndCapture(s, &g));
      SVM_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      SVM_HIP(hipGraphDestroy(g));
    }
    SVM_HIP(hipGraphLaunch(ge, s));
  } else {
    bk::launch_pass_clear(d_args, s);
    bk::launch_profile_masks(d_margs, s);
    if (nchk) bk::launch_drop_check(d_args, s);
    if (need_bits) bk::launch_slot_bits(d_args, s);
    SVM_HIP(hipEventRecord(eb, s));
    bk::launch_shared_scan(d_args, ch.variant, cap, grid, s);
    SVM_HIP(hipEventRecord(es, s));
    bk::launch_emit_count(d_emit, d_args, counts, offsets, totals, s);
  }
  std::vector<unsigned long long> tot(nq);
  unsigned long long dq = 0;
  SVM_HIP(hipMemcpyAsync(tot.data(), totals, nq * 8, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipMemcpyAsync(&dq, dropq, 8, hipMemcpyDeviceToHost, s));
  SVM_HIP(hipStreamSynchronize(s));
  // emission of rows (sized by the counts just read)
  uint64_t rows = 0;
  for (int i = 0; i < nq; ++i) {
    em[i].out0 = rows;
    rows += tot[i];
  }
  SVM_HIP(hipMemcpyAsync(dd + PassLayout::o_emit, em, nq * sizeof(bk::Emit), hipMemcpyHostToDevice, s));
  int32_t* keys = (int32_t*)sc.get(8, rows * 4 * bk::kMaxKey + 16);
  int64_t* vals = (int64_t*)sc.get(9, rows * 8 + 8);
  bk::launch_emit_rows(d_emit, nq, blocks, offsets, keys, vals, rows, s);
  std::vector<int32_t> hk(rows * bk::kMaxKey + 1);
  std::vector<int64_t> hv(rows + 1);
  if (rows) {
    SVM_HIP(hipMemcpyAsync(hk.data(), keys, rows * 4 * bk::kMaxKey, hipMemcpyDeviceToHost, s));
    SVM_HIP(hipMemcpyAsync(hv.data(), vals, rows * 8, hipMemcpyDeviceToHost, s));
  }
  SVM_HIP(hipEventRecord(e1, s));
  SVM_HIP(hipStreamSynchronize(s));
  float ms_all = 0;
  SVM_HIP(hipEventElapsedTime(&ms_all, e0, e1));
  if (E.cfg.verbose) {
    uint64_t slice_rows = 0;
    for (int q = 0; q < nq; ++q) slice_rows += (uint64_t)(qrow[2 * q + 1] - qrow[2 * q]);
    std::string lds_maps;
    for (int k = 0; k < A->nslot; ++k) {
      const bk::Slot& S = A->slot[k];
      lds_maps += k ? "," : "";
      lds_maps += S.lmap >= 0 ? "map" + std::to_string(1 << S.lmap_log2) : S.lbits >= 0 ? "bits" : "-";
    }
    char phases[128] = "";
    if (!graph) {
      float mb = 0, msc = 0;
      SVM_HIP(hipEventElapsedTime(&mb, e0, eb));
      SVM_HIP(hipEventElapsedTime(&msc, eb, es));
      snprintf(phases, sizeof(phases), ", masks %.3f, scan %.3f", mb, msc);
    }
    fprintf(stderr, "[svm] shared scan of %d queries: gpu %.3f ms%s (%s; %u profiles in %zu slots; %llu rows / %llu "
                    "slice rows; %d threads, %d LDS slots (%s), %s lookups, key bits %d, tile %lld, %u tiles, grid %d; "
                    "LDS maps [%s]; %d slots checked for dropped inserts, %d queries drop)\n",
            nq, ms_all, phases, graph ? "graph" : "eager", nprof, P.slots.size(), (unsigned long long)union_rows,
            (unsigned long long)slice_rows, bk::scan_variant_block(ch.variant), bk::scan_variant_hash_slots(ch.variant),
            bk::scan_variant_dense(ch.variant) ? "dense" : "hash",
            bk::scan_variant_parallel(ch.variant) ? "parallel" : "gated", ch.bits, (long long)ch.tile, A->ntiles, grid,
            lds_maps.c_str(), nchk, __builtin_popcountll(dq));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(eb);
  (void)hipEventDestroy(es);
  for (int q = 0; q < nq; ++q) {
    const int i = P.qi[q];
    if ((dq >> q) & 1ull) {
      redo.push_back(i);  // its reference hash table drops an insert: the plain path is exact
      continue;
    }
    auto r = std::make_unique<svm_result>();
    const int K = qs[i]->K;
    r->kind = K == 0 ? 2 : 1;
    r->rows = (int64_t)tot[q];
    for (int k = 0; k < K; ++k) {
      r->types.push_back({Int32, 4});
      std::vector<uint8_t> col(tot[q] * 4);
      if (tot[q]) memcpy(col.data(), hk.data() + (size_t)k * rows + em[q].out0, tot[q] * 4);
      r->cols.push_back(std::move(col));
    }
    r->types.push_back({Int64, 8});
    std::vector<uint8_t> col(tot[q] * 8);
    if (tot[q]) memcpy(col.data(), hv.data() + em[q].out0, tot[q] * 8);
    r->cols.push_back(std::move(col));
    r->scan_ms = ms_all;  // the whole pass completes together
    *outs[i] = std::move(r);
  }
}

// Fused execution of lowered queries that share one base fact table: greedy
// passes within the limits of one shared scan; a pass whose profiles exceed the
// LDS staging budget is split in two.  Queries whose dimension objects the
// reference's hash table would drop rows of get no result here: their indices (into
// qs) are appended to `redo` for the plain path.
static void run_fused(std::vector<FactLowered*>& qs, std::vector<std::unique_ptr<svm_result>*>& outs,
                      Engine& E, hipStream_t s, Scratch& sc, bool tune, std::vector<int>& redo) {
  if (qs.empty()) return;
  std::vector<PassB> passes;
  PassB cur;
  for (int i = 0; i < (int)qs.size(); ++i) {
    if (pass_add(cur, *qs[i], i, s)) continue;
    passes.push_back(std::move(cur));
    cur = PassB();
    if (!pass_add(cur, *qs[i], i, s)) fail(SVM_ERR_INVALID, "a lowered query exceeds the shared scan limits");
  }
  if (!cur.qi.empty()) passes.push_back(std::move(cur));
  while (!passes.empty()) {
    PassB P = std::move(passes.back());
    passes.pop_back();
    uint64_t nprof = 0;
    for (auto& sl : P.slots) {
      std::vector<int> attrs(sl.attrs.begin(), sl.attrs.end());
      auto pr = profile_of(sl.dim, sl.kcol, attrs, s);
      nprof += pr->ok ? pr->P + 1 : (uint64_t)bk::kMaxProfiles + 1;
    }
    if (nprof > (uint64_t)bk::kMaxProfiles && P.qi.size() > 1) {
      const size_t h = P.qi.size() / 2;
      passes.push_back(pass_of(qs, std::vector<int>(P.qi.begin(), P.qi.begin() + h), s));
      passes.push_back(pass_of(qs, std::vector<int>(P.qi.begin() + h, P.qi.end()), s));
      continue;
    }
    run_pass(qs, P, outs, E, s, sc, tune, redo);
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
  // plans decode independently: on the host workers
  parallel_for((size_t)n, [&](size_t i) {
    jobs[i].info = decode_pipeline_info(plans[i], lens[i]);
    jobs[i].plan = decode_plan(jobs[i].info);
  });
  const double t_decoded = host_ms();
  // dimension candidates, by pipeline id
  std::map<std::string, DimLowered> dims;
  for (auto& j : jobs) {
    DimLowered d = lower_dim(j.info, j.plan, E);
    if (d.ok) dims[j.info.id] = d;
  }
  const double t_dims = host_ms();
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
  std::map<std::string, int> job_of;
  for (int i = 0; i < n; ++i) job_of[jobs[i].info.id] = i;
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
    std::vector<int> redo;
    run_fused(fq, fo, E, s, *scratch, true, redo);
    for (int pos : redo) {
      // its dimension pipelines, then the query, on the plain path; the dimension
      // objects stay unmaterialised afterwards, like every absorbed one
      const int i = gq.second[pos];
      RunOptions dopt;
      dopt.want_host = false;
      for (auto& id : facts[i].deferred) run_pipeline(jobs[job_of.at(id)].info, dopt);
      RunOptions opt;
      opt.want_host = outs != nullptr;
      res[i] = run_pipeline(jobs[i].info, opt);
      for (auto& id : facts[i].deferred) E.store.drop(id);
      fused[i] = false;
    }
  }
  int nfused = 0;
  for (int i = 0; i < n; ++i) nfused += fused[i] ? 1 : 0;
  if (outs)
    for (int i = 0; i < n; ++i) outs[i] = res[i] ? res[i].release() : nullptr;
  if (E.cfg.verbose)
    fprintf(stderr, "[svm] batch_run %d plans, %d fused: host lowering %.3f ms (decode %.3f, dimensions %.3f, facts "
                    "%.3f), plain plans %.3f ms, total %.3f ms\n",
            n, nfused, t_lowered, t_decoded, t_dims - t_decoded, t_lowered - t_dims, t_plain - t_lowered, host_ms());
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
  uint64_t rows = 0;   // fact rows the query scans
  bool light = true;   // served in the light lane (not is_heavy)
  bool done = false;
  int status = SVM_OK;
  std::string error;
  std::unique_ptr<svm_result> result;
  std::chrono::steady_clock::time_point submitted, admitted, finished;
};

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  // waiting queries: the light lane (small fact slices) and the heavy lane (large
  // scans); in mode 1 every query waits in `light`, served in arrival order
  std::deque<std::shared_ptr<Ticket>> light, heavy;
  int nworkers = 1;
  std::map<uint64_t, std::shared_ptr<Ticket>> tickets;  // submitted and not yet claimed by a wait
  uint64_t bytes_in_use = 0, budget = 0, running = 0, completed = 0;
  int max_batch = 64, mode = 0;
  int submitting = 0;  // submissions between their stop check and their enqueue
  bool stop = false;
  std::vector<std::thread> workers;
  hipStream_t submit_stream = nullptr;
};

// The running server, and stopped servers that still hold unclaimed tickets (a
// wait on one of those returns the ticket's final status).  Every call copies the
// pointer it uses under g_server_mu, so a server stays alive while any call on it
// (a blocked wait, a submission) is in progress, whatever svm_server_stop does.
std::shared_ptr<Server> g_server;
std::vector<std::shared_ptr<Server>> g_stopped;
std::mutex g_server_mu;
// Ticket numbers are unique in the process, across servers.
std::atomic<uint64_t> g_next_ticket{1};
constexpr uint64_t kPlainNeed = 64ull << 20;  // HBM assumed for a plain query whose plans say too little
// A query is heavy when it scans more than half of its fact table and at least
// kHeavyMinRows rows (the serving mix's whole-table star joins); heavy queries wait in
// their own lane, so light ones never queue behind a whole-table pass.
constexpr uint64_t kHeavyMinRows = 4ull << 20;
bool is_heavy(uint64_t rows, uint64_t table_rows) { return rows >= kHeavyMinRows && rows > table_rows / 2; }
// Light queries waiting from which a heavy pass takes them along.
constexpr size_t kLightBacklog = 4;

// Seconds per fact row of a plain query's fact pipeline (map builds, scan, emission),
// learned from the served queries that scanned at least kRateMinRows rows, where the
// per-row part dominates (an exponential average; 3.7 ps per row, the rate of a scan
// that reads every column, until one has run).  The staged scans read far fewer bytes
// on selective queries, so the rate depends on the queries served.
constexpr uint64_t kRateMinRows = 64ull << 20;
std::mutex g_rate_mu;
double g_plain_rate = 3.7e-12;
double plain_rate() {
  std::lock_guard<std::mutex> g(g_rate_mu);
  return g_plain_rate;
}
void learn_plain_rate(const svm_result& r, uint64_t rows) {
  if (rows < kRateMinRows) return;
  const double sec = (r.build_ms + r.scan_ms + r.finish_ms) * 1e-3;
  if (!(sec > 0)) return;
  std::lock_guard<std::mutex> g(g_rate_mu);
  g_plain_rate = 0.75 * g_plain_rate + 0.25 * (sec / (double)rows);
}

// Whether one shared scan over `batch` beats running its queries one by one on the
// plain kernels, by a cost model fitted to measured passes on the 600M-row tables:
// a pass costs ~4.1 ps per row of the union of the slices (the issue-bound shared
// scan) + ~0.45 ps per slice row (each query's share) + ~0.2 ms (profile masks,
// emission, synchronisation); a plain query plain_rate() per slice row (its fact
// pipeline, as measured on the queries served) + ~0.15 ms (its dimension selects).  Queries
// whose slices barely overlap are cheaper alone.  A fused pass answers all its
// queries at its end, while queries run one by one finish in turn, so a pass is
// chosen only for a clear gain (under 0.7 of the plain time): latency stays the plain
// kernels' until sharing the scan saves real work.
bool fusing_pays(const std::vector<std::shared_ptr<Ticket>>& batch) {
  if (batch.size() < 2) return false;
  std::vector<std::pair<long long, long long>> r;
  double sum = 0;
  for (auto& t : batch) {
    r.push_back({t->fact.row0, t->fact.row1});
    sum += (double)(t->fact.row1 - t->fact.row0);
  }
  std::sort(r.begin(), r.end());
  double uni = 0;
  long long lo = r[0].first, hi = r[0].second;
  for (auto& x : r) {
    if (x.first > hi) {
      uni += (double)(hi - lo);
      lo = x.first;
      hi = x.second;
    } else {
      hi = std::max(hi, x.second);
    }
  }
  uni += (double)(hi - lo);
  const double fused = 4.1e-12 * uni + 0.45e-12 * sum + 0.2e-3;
  const double plain = plain_rate() * sum + 0.15e-3 * (double)batch.size();
  return fused < 0.7 * plain;
}

// HBM a query on the plain path holds while it runs, estimated from its plans and
// the resident tables' cached key ranges and dictionaries: each dimension pipeline's
// output (its source's rows x the projected widths), each join's dense map (4 bytes
// per key of the dimension key column's range, plus existence bits; a key column too
// wide for a map probes the reference's hash table, 5 slots x 4 bytes per row), and
// the group table (8 bytes per group and metric over the product of the group key
// columns' distinct values, plus flag bits; beyond 2^27 groups the sorted entries,
// 8 + 8 x metrics bytes per row for at most 2^24 rows before they are resized).
uint64_t plain_need(const std::vector<Job>& jobs, Engine& E, hipStream_t s) {
  try {
    std::map<std::string, size_t> made;
    for (size_t i = 0; i + 1 < jobs.size(); ++i) made[jobs[i].info.id] = i;
    uint64_t need = 4096;
    // the column a ref object's column `col` holds: (resident table, its column)
    auto origin = [&](const std::string& id, int col) -> std::pair<TablePtr, int> {
      auto it = made.find(id);
      if (it == made.end()) return {E.store.find(id), col};
      const Job& d = jobs[it->second];
      if (d.info.pre_pipeline_id.empty() || col < 0 || col >= (int)d.plan.select_exp.size()) return {nullptr, -1};
      int obj = -1, c = -1;
      if (!parse_operand(d.plan.select_exp[col], obj, c) || obj != 0) return {nullptr, -1};
      return {E.store.find(d.info.pre_pipeline_id[0]), c};
    };
    for (size_t i = 0; i + 1 < jobs.size(); ++i) {
      TablePtr src = jobs[i].info.pre_pipeline_id.empty() ? nullptr : E.store.find(jobs[i].info.pre_pipeline_id[0]);
      uint64_t width = 0;
      for (auto& c : jobs[i].plan.output_info.column_list) width += (uint64_t)std::max<int64_t>(c.val_size, 4);
      need += (src ? src->rows : 0) * width;
    }
    const Job& f = jobs.back();
    TablePtr fsrc = f.info.pre_pipeline_id.empty() ? nullptr : E.store.find(f.info.pre_pipeline_id[0]);
    if (!fsrc) return kPlainNeed;
    const uint64_t rows = fsrc->rows;
    for (size_t j = 0; j < f.plan.join_rkey.size(); ++j) {
      const KeyIndex R = f.plan.join_rkey[j];
      if (R.obj < 1 || R.obj >= (int)f.info.pre_pipeline_id.size()) return kPlainNeed;
      auto o = origin(f.info.pre_pipeline_id[R.obj], R.col);
      if (!o.first || o.second < 0 || o.second >= (int)o.first->cols.size()) return kPlainNeed;
      uint64_t range = 0;
      if (o.first->cols[o.second].type == Int32 && o.first->rows) {
        KeyEntry ke = table_key(o.first, o.second, s);
        range = (uint64_t)((int64_t)ke.kmax - (int64_t)ke.kmin + 1);
      }
      need += range && range <= (1ull << 28) ? range * 4 + range / 8 : o.first->rows * 20;
    }
    uint64_t G = 1;
    for (const KeyIndex& k : f.plan.reduce_key) {
      uint64_t card = 0;
      if (k.obj == 0) {
        if (k.col < 0 || k.col >= (int)fsrc->cols.size() || fsrc->cols[k.col].type != Int32) return kPlainNeed;
        KeyEntry ke = table_key(fsrc, k.col, s);
        card = (uint64_t)((int64_t)ke.kmax - (int64_t)ke.kmin + 1);
      } else {
        if (k.obj >= (int)f.info.pre_pipeline_id.size()) return kPlainNeed;
        auto o = origin(f.info.pre_pipeline_id[k.obj], k.col);
        if (!o.first || o.second < 0 || o.second >= (int)o.first->cols.size()) return kPlainNeed;
        card = table_dict(o.first, o.second, s)->n;
      }
      G = card && G > (1ull << 40) / card ? (1ull << 40) : G * std::max<uint64_t>(card, 1);
    }
    const uint64_t M = std::max<size_t>(f.plan.metric_exp.size(), 1);
    if (G <= (1ull << 27)) need += G * M * 8 + G / 8 + 64;
    else need += std::min<uint64_t>(rows, 1ull << 24) * (8 + 8 * M);
    return need;
  } catch (const Error&) {
    return kPlainNeed;
  }
}

std::shared_ptr<Server> running_server() {
  std::lock_guard<std::mutex> g(g_server_mu);
  return g_server;
}

// The server (running or stopped) that holds `ticket`, or null.
std::shared_ptr<Server> server_of(uint64_t ticket) {
  std::lock_guard<std::mutex> g(g_server_mu);
  auto holds = [&](const std::shared_ptr<Server>& S) {
    std::lock_guard<std::mutex> lk(S->mu);
    return S->tickets.count(ticket) > 0;
  };
  if (g_server && holds(g_server)) return g_server;
  for (auto& S : g_stopped)
    if (holds(S)) return S;
  return nullptr;
}

// A served query's pipelines write their outputs into the shared store under their
// plan ids.  Two queries in flight may carry the same ids (the reference's drivers
// use fixed ones, e.g. test/test-SSB-Q21.cu:21), so every id a ticket's pipelines
// produce is renamed to one private to the ticket, in the pipelines' own ids and in
// the pre_pipeline_id lists that name them; objects the query reads but does not
// produce (its fact table and dimension tables) keep their ids.
void namespace_ticket(std::vector<Job>& jobs, uint64_t tag) {
  std::map<std::string, std::string> ren;
  const std::string suffix = std::string("\x01ticket") + std::to_string(tag);
  for (auto& j : jobs) ren[j.info.id] = j.info.id + suffix;
  for (auto& j : jobs) {
    j.info.id = ren[j.info.id];
    for (auto& pre : j.info.pre_pipeline_id) {
      auto it = ren.find(pre);
      if (it != ren.end()) pre = it->second;
    }
  }
}

// A served query is self-contained (its dimension pipelines, then its fact
// pipeline): the objects it materialises are its intermediates and are dropped
// once its result is on the host, so a long-running server does not accumulate them.
void run_plain_query(Ticket& t, hipStream_t s) {
  std::unique_ptr<svm_result> last;
  Engine& E = engine();
  try {
    for (auto& j : t.jobs) {
      RunOptions opt;
      opt.stream = s;
      opt.want_host = true;
      last = run_pipeline(j.info, opt);
    }
  } catch (...) {
    for (auto& j : t.jobs) E.store.drop(j.info.id);
    throw;
  }
  if (last) learn_plain_rate(*last, t.rows);
  t.result = std::move(last);
  for (auto& j : t.jobs) E.store.drop(j.info.id);
}

void worker_main(Server* S, int index) {
  Engine& E = engine();
  hipStream_t s = nullptr;
  SVM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  Scratch sc;
  for (;;) {
    std::vector<std::shared_ptr<Ticket>> batch;
    uint64_t need = 0;
    bool alone = true;
    {
      // Worker 0 of a batching server with two or more workers serves the light lane
      // only, so light queries never wait behind a heavy pass on every worker; the
      // others take heavy queries first.
      const bool light_only = S->mode == 0 && S->nworkers >= 2 && index == 0;
      auto fits = [&](const std::deque<std::shared_ptr<Ticket>>& q) {
        return !q.empty() && S->bytes_in_use + q.front()->need <= S->budget;
      };
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || fits(S->light) || (!light_only && fits(S->heavy)); });
      if (S->stop) break;
      const bool heavy_lane = !light_only && fits(S->heavy);
      // candidates in arrival order while memory and the batch size allow (a batch is
      // one shared scan over one fact table); they run as one pass when sharing the
      // scan pays, else the oldest runs alone on the plain kernels and the rest wait
      // for the next free worker
      const int cap = S->mode == 1 ? 1 : S->max_batch;
      auto take = [&](std::deque<std::shared_ptr<Ticket>>& q) {
        while (!q.empty() && (int)batch.size() < cap && S->bytes_in_use + need + q.front()->need <= S->budget) {
          auto t = q.front();
          if (!batch.empty() && (!t->fused || !batch[0]->fused || t->fact.base != batch[0]->fact.base)) break;
          need += t->need;
          batch.push_back(t);
          q.pop_front();
          if (!t->fused) break;
        }
      };
      take(heavy_lane ? S->heavy : S->light);
      // a heavy pass reads the whole table anyway: when the light lane has a backlog,
      // waiting light queries ride along for their share of the pass
      if (heavy_lane && S->mode == 0 && batch[0]->fused && S->light.size() >= kLightBacklog) take(S->light);
      alone = S->mode == 1 || !batch[0]->fused || !fusing_pays(batch);
      if (alone && batch.size() > 1) {
        // back to the front of their lanes, in arrival order
        while (batch.size() > 1) {
          need -= batch.back()->need;
          (batch.back()->light ? S->light : S->heavy).push_front(batch.back());
          batch.pop_back();
        }
        // the queries put back are another worker's to take (every worker re-checks: a
        // light-lane worker cannot take a heavy query)
        S->cv_work.notify_all();
      }
      S->bytes_in_use += need;
      S->running += batch.size();
      const auto now = std::chrono::steady_clock::now();
      for (auto& t : batch) t->admitted = now;
    }
    const auto t_run0 = std::chrono::steady_clock::now();
    try {
      if (alone) {
        run_plain_query(*batch[0], s);
      } else {
        std::vector<FactLowered*> fq;
        std::vector<std::unique_ptr<svm_result>*> fo;
        for (auto& t : batch) {
          fq.push_back(&t->fact);
          fo.push_back(&t->result);
        }
        // shapes not tuned by an earlier batch_run take the default launch shape:
        // a serving batch does not stall on tuning trials
        std::vector<int> redo;
        run_fused(fq, fo, E, s, sc, false, redo);
        for (int pos : redo) run_plain_query(*batch[pos], s);
      }
    } catch (const Error& e) {
      for (auto& t : batch) { t->status = e.code; t->error = e.what(); }
    } catch (const std::exception& e) {
      for (auto& t : batch) { t->status = SVM_ERR_DEVICE; t->error = e.what(); }
    }
    if (E.cfg.verbose) {
      int light = 0;
      uint64_t rows = 0;
      for (auto& t : batch) {
        light += t->light ? 1 : 0;
        rows += t->rows;
      }
      fprintf(stderr, "[svm] server worker %d: %s of %zu queries (%d light), %llu slice rows, %.3f ms\n", index,
              alone ? "plain run" : "shared scan", batch.size(), light, (unsigned long long)rows,
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_run0).count());
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->bytes_in_use -= need;
      S->running -= batch.size();
      S->completed += batch.size();
      const auto now = std::chrono::steady_clock::now();
      for (auto& t : batch) {
        t->finished = now;
        t->done = true;
      }
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
  auto S = std::make_shared<Server>();
  S->budget = budget ? budget : E.cfg.hbm_budget;
  S->max_batch = max_batch > 0 ? max_batch : 64;
  S->mode = mode;
  S->nworkers = workers;
  SVM_HIP(hipStreamCreateWithFlags(&S->submit_stream, hipStreamNonBlocking));
  Server* raw = S.get();
  for (int w = 0; w < workers; ++w) S->workers.emplace_back(worker_main, raw, w);
  g_server = std::move(S);
  return SVM_OK;
}

int server_submit(const uint8_t* const* plans, const size_t* lens, int n, uint64_t* ticket) {
  std::shared_ptr<Server> S = running_server();
  if (!S) fail(SVM_ERR_STATE, "server not running");
  if (n <= 0 || !plans || !lens) fail(SVM_ERR_INVALID, "a query needs at least its fact pipeline");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) fail(SVM_ERR_STATE, "server stopping");
    S->submitting++;  // svm_server_stop waits for this submission (it uses submit_stream)
  }
  struct Leave {
    Server* S;
    ~Leave() {
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->submitting--;
      }
      S->cv_done.notify_all();
    }
  } leave{S.get()};
  Engine& E = engine();
  auto t = std::make_shared<Ticket>();
  t->id = g_next_ticket++;
  t->jobs.resize(n);
  for (int i = 0; i < n; ++i) t->jobs[i].info = decode_pipeline_info(plans[i], lens[i]);
  namespace_ticket(t->jobs, t->id);
  for (int i = 0; i < n; ++i) t->jobs[i].plan = decode_plan(t->jobs[i].info);
  if (S->mode == 0) {
    std::map<std::string, DimLowered> dims;
    for (int i = 0; i + 1 < n; ++i) {
      DimLowered d = lower_dim(t->jobs[i].info, t->jobs[i].plan, E);
      if (d.ok) dims[t->jobs[i].info.id] = d;
    }
    t->fact = lower_fact(t->jobs[n - 1], dims, E, S->submit_stream);
    // every dimension pipeline of the query must be absorbed by the fused fact pipeline
    t->fused = t->fact.ok && t->fact.deferred.size() == (size_t)(n - 1);
  }
  uint64_t table_rows = 0;
  if (t->fused) {
    t->rows = (uint64_t)(t->fact.row1 - t->fact.row0);
    table_rows = t->fact.base ? t->fact.base->rows : t->rows;
  } else if (!t->jobs.back().info.pre_pipeline_id.empty()) {
    TablePtr src = E.store.find(t->jobs.back().info.pre_pipeline_id[0]);
    t->rows = src ? src->rows : 0;
    table_rows = src ? (src->base ? src->base->rows : src->rows) : 0;
  }
  t->light = S->mode == 1 || !is_heavy(t->rows, table_rows);
  t->need = t->fused ? fused_bytes(t->fact) : plain_need(t->jobs, E, S->submit_stream);
  if (t->need > S->budget) fail(SVM_ERR_BUSY, "query needs more HBM than the serving budget");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) fail(SVM_ERR_STATE, "server stopping");
    t->submitted = std::chrono::steady_clock::now();
    S->tickets[t->id] = t;
    (t->light ? S->light : S->heavy).push_back(t);
    if (ticket) *ticket = t->id;
  }
  // every worker re-checks its lanes: waking one that cannot serve this query's lane
  // (the light-only worker, for a heavy query) would leave the query waiting
  S->cv_work.notify_all();
  return SVM_OK;
}

int server_wait(uint64_t ticket, int timeout_ms, svm_result** out, double* queued_ms, double* total_ms) {
  std::shared_ptr<Server> S = server_of(ticket);
  if (!S) {
    if (!running_server() && ticket < g_next_ticket.load()) fail(SVM_ERR_STATE, "server not running");
    fail(SVM_ERR_NOT_FOUND, "unknown ticket " + std::to_string(ticket));
  }
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) fail(SVM_ERR_NOT_FOUND, "unknown ticket " + std::to_string(ticket));
  auto t = it->second;
  auto ready = [&] { return t->done; };
  if (timeout_ms < 0) S->cv_done.wait(lk, ready);
  else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
    return SVM_ERR_BUSY;
  if (!S->tickets.erase(ticket)) fail(SVM_ERR_NOT_FOUND, "ticket " + std::to_string(ticket) + " already claimed");
  const bool drained = S->stop && S->tickets.empty();
  lk.unlock();
  if (drained) {
    // a stopped server whose last ticket is claimed is forgotten
    std::lock_guard<std::mutex> g(g_server_mu);
    g_stopped.erase(std::remove(g_stopped.begin(), g_stopped.end(), S), g_stopped.end());
  }
  auto ms = [&](std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double, std::milli>(a - t->submitted).count();
  };
  if (queued_ms) *queued_ms = t->admitted.time_since_epoch().count() ? ms(t->admitted) : 0.0;
  if (total_ms) *total_ms = ms(t->finished);
  if (t->status != SVM_OK) fail(t->status, t->error);
  if (out) *out = t->result.release();
  return SVM_OK;
}

int server_stats(uint64_t* queued, uint64_t* running, uint64_t* done, uint64_t* bytes) {
  std::shared_ptr<Server> S = running_server();
  if (!S) fail(SVM_ERR_STATE, "server not running");
  std::lock_guard<std::mutex> lk(S->mu);
  if (queued) *queued = S->light.size() + S->heavy.size();
  if (running) *running = S->running;
  if (done) *done = S->completed;
  if (bytes) *bytes = S->bytes_in_use;
  return SVM_OK;
}

int server_stop() {
  std::shared_ptr<Server> S;
  {
    std::lock_guard<std::mutex> g(g_server_mu);
    S = std::move(g_server);
    g_server.reset();
  }
  if (!S) return SVM_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  // admitted batches finish; workers take nothing new
  for (auto& w : S->workers) w.join();
  S->workers.clear();
  {
    // queries never admitted fail with SVM_ERR_STATE; submissions in progress end
    // before the submission stream goes away
    std::unique_lock<std::mutex> lk(S->mu);
    const auto now = std::chrono::steady_clock::now();
    for (auto* q : {&S->light, &S->heavy}) {
      for (auto& t : *q) {
        t->status = SVM_ERR_STATE;
        t->error = "server stopped before the query was admitted";
        t->finished = now;
        t->done = true;
      }
      q->clear();
    }
    S->cv_done.notify_all();
    S->cv_done.wait(lk, [&] { return S->submitting == 0; });
  }
  (void)hipStreamDestroy(S->submit_stream);
  S->submit_stream = nullptr;
  std::lock_guard<std::mutex> g(g_server_mu);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->tickets.empty()) g_stopped.push_back(S);
  }
  return SVM_OK;
}

}  // namespace svm
