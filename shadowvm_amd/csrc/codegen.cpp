// codegen.cpp — emits the HIP source of the fused per-pipeline kernels.
//
// Where the reference emits one thread per row that probes a host-built Hash31
// table and writes unreduced key/metric rows for a host sort (codegen/template.cpp:77-89,
// codegen/codegen.cpp:100-229), the kernels here are persistent grid-stride scans
// over HBM-resident int32 columns:
//   * 128-bit non-temporal loads of 4 rows per column per thread;
//   * filter predicates inlined as C int expressions (the reference language);
//   * each join a lookup in a bit-packed dense map (optionally staged in LDS)
//     whose slot holds existence, the group code of the dimension row, or its row id;
//   * group-by into a per-workgroup LDS table (dense array, or open-addressing hash
//     for large group spaces) merged into HBM with 64-bit atomics at the end;
//   * select pipelines compact passing rows in row order with wave64 ballots.
#include "codegen.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>

#include "kargs.h"

namespace svm {

uint64_t source_hash(const std::string& s);

// The runtime the generated code calls (Hash / StrHash / StrCmp / StrCopy / Probe of
// com/cpp/cpu-runtime.hpp:37-152 and the packed-map lookup).  The same text is
// compiled into the device kernels (SVM_FN = __device__ __forceinline__) and into
// the host engine's code (SVM_FN = static inline), so both evaluate alike.
static const char* kCommonRT = R"SVMRT(
typedef unsigned long long svm_u64;
typedef long long svm_i64;

/* Hash31 over n bytes read as signed char, uint32 wrap, top bit cleared
   (reference com/cpp/cuda-runtime.cu:64-71). */
SVM_FN unsigned Hash(const char* p, int n) {
  unsigned h = 0;
  for (int k = 0; k < n; ++k) h = h * 31u + (unsigned)(int)(signed char)p[k];
  return h & 0x7FFFFFFFu;
}
SVM_FN unsigned StrHash(const char* p, int n) {
  unsigned h = 0;
  for (int k = 0; k < n && p[k] != '\0'; ++k) h = h * 31u + (unsigned)(int)(signed char)p[k];
  return h & 0x7FFFFFFFu;
}
)SVMRT";

// StrCmp, prefix equality (com/cpp/cuda-runtime.cu:34-54): 0 once either side
// reaches NUL, else the first byte difference.  Every operand is a row of a padded
// column or a zero-tailed literal, so up to 64 bytes (the widths of fixed string
// columns) both sides are read in full before the answer is selected without
// branches, and the result is the same as the early-exit loop's.  The host engine
// reads them byte by byte; the kernels read whole aligned 32-bit words and take each
// operand's bytes out of word pairs (v_alignbyte), a quarter of the loads.
static const char* kHostStrCmp = R"SVMRT(
SVM_FN int StrCmp(const char* a, const char* b, int n) {
  if (n <= 64) {
    int r = 0;
    bool done = false;
    #pragma unroll
    for (int k = 0; k < 64; ++k) {
      if (k >= n) break;
      const char x = a[k], y = b[k];
      const int d = (int)(signed char)x - (int)(signed char)y;
      const bool stop = (x == '\0') | (y == '\0');
      r = (!done && !stop && d != 0) ? d : r;
      done = done | stop | (d != 0);
    }
    return r;
  }
  for (int k = 0; k < n; ++k) {
    char x = a[k], y = b[k];
    if (x == '\0' || y == '\0') return 0;
    if (x != y) return (int)(signed char)x - (int)(signed char)y;
  }
  return 0;
}
)SVMRT";

static const char* kDeviceStrCmp = R"SVMRT(
SVM_FN int StrCmp(const char* a, const char* b, int n) {
  if (n <= 64) {
    const svm_u64 pa = (svm_u64)a, pb = (svm_u64)b;
    const unsigned* wa = (const unsigned*)(pa & ~3ull);
    const unsigned* wb = (const unsigned*)(pb & ~3ull);
    const unsigned sa = (unsigned)(pa & 3ull), sb = (unsigned)(pb & 3ull);
    int r = 0;
    bool done = false;
    unsigned ca = wa[0], cb = wb[0];
    #pragma unroll
    for (int m = 0; m < 16; ++m) {
      if (4 * m >= n) break;
      const unsigned na = wa[m + 1], nb = wb[m + 1];
      const unsigned xa = __builtin_amdgcn_alignbyte(na, ca, sa), xb = __builtin_amdgcn_alignbyte(nb, cb, sb);
      ca = na;
      cb = nb;
      #pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (4 * m + j >= n) break;
        const int x = (int)(signed char)(xa >> (8 * j)), y = (int)(signed char)(xb >> (8 * j));
        const int d = x - y;
        const bool stop = (x == 0) | (y == 0);
        r = (!done && !stop && d != 0) ? d : r;
        done = done | stop | (d != 0);
      }
    }
    return r;
  }
  for (int k = 0; k < n; ++k) {
    char x = a[k], y = b[k];
    if (x == '\0' || y == '\0') return 0;
    if (x != y) return (int)(signed char)x - (int)(signed char)y;
  }
  return 0;
}
)SVMRT";

static const char* kCommonRT2 = R"SVMRT(
SVM_FN void StrCopy(char* d, const char* s, int n) {
  for (int k = 0; k < n; ++k) d[k] = s[k];
}

template <int W>
SVM_FN unsigned svm_map_get(const unsigned* m, unsigned off) {
  if (W == 32) return m[off];
  constexpr int E = 32 / W;
  constexpr int S = W == 1 ? 5 : W == 2 ? 4 : W == 4 ? 3 : W == 8 ? 2 : 1;
  return (m[off >> S] >> ((off & (unsigned)(E - 1)) * (unsigned)W)) & ((1u << W) - 1u);
}

/* Probe<T> (reference com/cpp/cpu-runtime.hpp:125-138): walk at most 10 slots from
   Hash(key bytes) % slots; -1 at an empty slot, the row whose key equals, -1 after 10. */
template <typename T>
SVM_FN int svm_probe(const T* key, int key_size, const T* tab, const int* ht, unsigned slots) {
  unsigned idx = Hash((const char*)key, key_size) % slots;
  for (int t = 0; t < 10; ++t) {
    const int pos = ht[idx];
    if (pos < 0) return -1;
    if (tab[pos] == *key) return pos;
    idx = idx + 1u == slots ? 0u : idx + 1u;
  }
  return -1;
}
/* Probe<char> (:140-152): StrHash home, StrCmp (prefix) equality, and row 0 (not -1)
   after 10 slots. */
SVM_FN int svm_probe_str(const char* key, int key_size, const char* tab, const int* ht,
                                             unsigned slots) {
  unsigned idx = StrHash(key, key_size) % slots;
  for (int t = 0; t < 10; ++t) {
    const int pos = ht[idx];
    if (pos < 0) return -1;
    if (StrCmp(key, tab + (long long)pos * key_size, key_size) == 0) return pos;
    idx = idx + 1u == slots ? 0u : idx + 1u;
  }
  return 0;
}

#define PI(k) (a.pi[k])
#define PS(off) ((const char*)(a.sblob + (off)))
)SVMRT";

// Device-only helpers of the generated kernels.
static const char* kDeviceOnlyRT = R"SVMRT(
typedef int svm_i32x4 __attribute__((ext_vector_type(4)));
typedef int svm_i32x2 __attribute__((ext_vector_type(2)));
/* Reserves one entry of an append buffer per active lane with `want` set: one atomic
   per wave on the counter, each lane's position its rank among the wanting lanes. */
__device__ __forceinline__ svm_u64 svm_wave_reserve(svm_u64* counter, bool want) {
  const svm_u64 m = __ballot(want);
  if (!m) return 0ull;
  const int lane = (int)__lane_id();
  const int leader = __ffsll((long long)m) - 1;
  svm_u64 base = 0ull;
  if (lane == leader) base = atomicAdd(counter, (svm_u64)__popcll(m));
  base = __shfl(base, leader);
  return base + (svm_u64)__popcll(m & ((1ull << lane) - 1ull));
}
/* StrCopy with the widest accesses both addresses allow: four bytes, two, or one at a
   time (the same bytes; a projection's rows are n-byte aligned in both columns). */
__device__ __forceinline__ void svm_str_copy(char* d, const char* s, int n) {
  const unsigned long long al = ((unsigned long long)d | (unsigned long long)s | (unsigned long long)n);
  if ((al & 3ull) == 0ull) {
    for (int k = 0; k < n; k += 4) *(unsigned*)(d + k) = *(const unsigned*)(s + k);
  } else if ((al & 1ull) == 0ull) {
    for (int k = 0; k < n; k += 2) *(unsigned short*)(d + k) = *(const unsigned short*)(s + k);
  } else {
    for (int k = 0; k < n; ++k) d[k] = s[k];
  }
}

)SVMRT";

std::string device_runtime_text() {
  return std::string("#define SVM_FN __device__ __forceinline__\n") + kCommonRT + kDeviceStrCmp + kCommonRT2 +
         kDeviceOnlyRT;
}
std::string host_runtime_text() {
  return std::string("#define SVM_FN static inline\n") + kCommonRT + kHostStrCmp + kCommonRT2;
}

namespace {

std::string I(int v) { return std::to_string(v); }

void emit_prologue(std::ostringstream& o, const KernelShape& s) {
  o << SVM_KARGS_TEXT << "\n" << device_runtime_text() << "\n";
  o << "#define SVM_BLOCK " << s.block << "\n";
}

// Column pointer macros used by lowered expressions.
void emit_colptrs(std::ostringstream& o, const KernelShape& s) {
  for (int c : s.fact_other)
    o << "  const char* FCOL" << c << " = (const char*)a.ptr[" << c << "];\n";
  for (auto& kv : s.dcol_slot)
    o << "  const char* DCOL" << kv.first.first << "_" << kv.first.second
      << " = (const char*)a.ptr[" << kv.second << "];\n";
}

const char* key_ctype(int t) { return t == 1 ? "short" : t == 3 ? "long long" : t == 0 ? "char" : "int"; }

// Statements computing `int <var>` = the dimension row a probe join finds for source
// row i (Probe<T> / Probe<char> on the reference's hash table in a.ptr[map slot],
// a.iv[4j+1] slots), as GenJoinBlock emits it (codegen/codegen.cpp:137-141): the key
// is read at the dimension key's width.
std::string probe_stmt(const JoinShape& J, size_t j, const std::string& var) {
  std::ostringstream o;
  const std::string ht = "(const int*)a.ptr[" + I(SVM_SLOT_MAP + (int)j) + "]";
  const std::string slots = "(unsigned)a.iv[" + I(4 * (int)j + 1) + "]";
  const std::string tab = "DCOL" + I(J.robj) + "_" + I(J.rcol);
  if (J.kind == JOIN_PROBE_STR) {
    o << "const int " << var << " = svm_probe_str((const char*)(FCOL" << J.lcol << " + i * " << J.ksize << "ll), "
      << J.ksize << ", " << tab << ", " << ht << ", " << slots << ");";
  } else if (J.ktype == 2) {  // Int32 keys are held in registers
    o << "const int " << var << "k = f" << J.lcol << "; const int " << var << " = svm_probe<int>(&" << var
      << "k, 4, (const int*)" << tab << ", " << ht << ", " << slots << ");";
  } else {
    const char* ct = key_ctype(J.ktype);
    o << "const int " << var << " = svm_probe<" << ct << ">((const " << ct << "*)(FCOL" << J.lcol << " + i * "
      << J.ksize << "ll), " << J.ksize << ", (const " << ct << "*)" << tab << ", " << ht << ", " << slots << ");";
  }
  return o.str();
}

void emit_map_ptrs(std::ostringstream& o, const KernelShape& s) {
  for (size_t j = 0; j < s.joins.size(); ++j) {
    const JoinShape& J = s.joins[j];
    if (J.lds)
      o << "  const unsigned* m" << j << " = svm_lds + a.iv[" << 4 * j + 3 << "];\n";
    else
      o << "  const unsigned* m" << j << " = (const unsigned*)a.ptr[" << SVM_SLOT_MAP + j << "];\n";
    // a deferred map's existence bits (after its packed words, or staged in LDS)
    if (J.kind == JOIN_MAP && J.defer) {
      if (J.bits_lds)
        o << "  const unsigned* mb" << j << " = svm_lds + a.iv[" << 4 * j + 3 << "];\n";
      else
        o << "  const unsigned* mb" << j << " = m" << j << " + a.iv[" << 4 * j + 2 << "];\n";
    }
  }
}

void emit_stage_maps(std::ostringstream& o, const KernelShape& s) {
  bool any = false;
  for (size_t j = 0; j < s.joins.size(); ++j) {
    if (s.joins[j].bits_lds) {
      // the existence bits after the packed map: range / 32 + 1 words (bit `range` is 0)
      o << "  { const unsigned* g = (const unsigned*)a.ptr[" << SVM_SLOT_MAP + j << "] + a.iv[" << 4 * j + 2 << "];\n"
        << "    unsigned* d = svm_lds + a.iv[" << 4 * j + 3 << "];\n"
        << "    const int nw = (int)(((unsigned)a.iv[" << 4 * j + 1 << "] >> 5) + 1u);\n"
        << "    for (int w = threadIdx.x; w < nw; w += SVM_BLOCK) d[w] = g[w]; }\n";
      continue;
    }
    if (!s.joins[j].lds) continue;
    any = true;
    o << "  { const unsigned* g = (const unsigned*)a.ptr[" << SVM_SLOT_MAP + j << "];\n"
      << "    unsigned* d = svm_lds + a.iv[" << 4 * j + 3 << "];\n"
      << "    for (int w = threadIdx.x; w < a.iv[" << 4 * j + 2 << "]; w += SVM_BLOCK) d[w] = g[w]; }\n";
  }
  (void)any;
}

// Filter, then every join in plan order (each guarded by the running condition,
// as GenJoinBlock guards Probe with if(cond), codegen/codegen.cpp:141-145).
void emit_row_body_joins(std::ostringstream& o, const KernelShape& s, const std::string& ind) {
  o << ind << "if (cond) cond = (" << s.filter << ");\n";
  for (size_t j = 0; j < s.joins.size(); ++j) {
    const JoinShape& J = s.joins[j];
    if (J.kind != JOIN_MAP) {
      // p = row + 1, so 0 still means "no row"
      o << ind << "unsigned p" << j << " = 0u;\n"
        << ind << "if (cond) {\n"
        << ind << "  " << probe_stmt(J, j, "r" + I((int)j)) << "\n"
        << ind << "  p" << j << " = (unsigned)(r" << j << " + 1);\n"
        << ind << "  cond = r" << j << " >= 0;\n"
        << ind << "}\n"
        << ind << "const long long JI" << J.robj << " = (long long)p" << j << " - 1;\n";
      continue;
    }
    if (J.defer) {
      // existence bit first; the payload only for keys that exist
      o << ind << "unsigned p" << j << " = 0u;\n"
        << ind << "if (cond) {\n"
        << ind << "  const unsigned off = (unsigned)(f" << J.lcol << " - a.iv[" << 4 * j << "]);\n"
        << ind << "  if (off < (unsigned)a.iv[" << 4 * j + 1 << "] && svm_map_get<1>(mb" << j
        << ", off)) p" << j << " = svm_map_get<" << J.wbits << ">(m" << j << ", off);\n"
        << ind << "  cond = p" << j << " != 0u;\n"
        << ind << "}\n";
      continue;
    }
    o << ind << "unsigned p" << j << " = 0u;\n"
      << ind << "if (cond) {\n"
      << ind << "  const unsigned off = (unsigned)(f" << J.lcol << " - a.iv[" << 4 * j << "]);\n"
      << ind << "  if (off < (unsigned)a.iv[" << 4 * j + 1 << "]) p" << j << " = svm_map_get<"
      << J.wbits << ">(m" << j << ", off);\n"
      << ind << "  cond = p" << j << " != 0u;\n"
      << ind << "}\n";
    if (J.payload == PAY_ROW)
      o << ind << "const long long JI" << J.robj << " = (long long)p" << j << " - 1;\n";
  }
}

// The group key of one passing row (`gk`, packed, for AGG_SORT; else the dense group
// id `gid`) and its metric values `mv<m>`, emitted inside `if (cond) {` with the
// row's f<c>, p<j> and JI<o> in scope (device kernels and host engine alike).
void emit_key_and_metrics(std::ostringstream& o, const KernelShape& s) {
  const int M = (int)s.metrics.size();
  if (s.agg == AGG_SORT) {
    // packed 64-bit key: each component's code at its bit offset
    o << "        svm_u64 gk = 0ull;\n";
    for (size_t q = 0; q < s.keys.size(); ++q) {
      const KeyShape& K = s.keys[q];
      std::string code;
      if (K.src == KEY_DIMCODE)
        code = "(p" + I(K.join) + " - 1u)";
      else if (K.src == KEY_DIMROW)
        code = "(unsigned)((const int*)a.ptr[" + I(SVM_SLOT_CODE + (int)q) + "])[JI" + I(K.obj) + "]";
      else
        code = "(unsigned)(f" + I(K.col) + " - a.iv[" + I(32 + 4 * (int)q) + "])";
      o << "        gk |= (svm_u64)" << code << " << a.iv[" << 32 + 4 * q + 3 << "];\n";
    }
  }
  // group id
  if (s.agg != AGG_SCALAR && s.agg != AGG_SORT) {
    o << "        unsigned gid = 0u;\n";
    for (size_t q = 0; q < s.keys.size(); ++q) {
      const KeyShape& K = s.keys[q];
      std::string stride = "(unsigned)a.iv[" + I(32 + 4 * (int)q + 1) + "]";
      if (K.src == KEY_DIMCODE)
        o << "        gid += (p" << K.join << " - 1u) * " << stride << ";\n";
      else if (K.src == KEY_DIMROW)
        o << "        gid += (unsigned)((const int*)a.ptr[" << SVM_SLOT_CODE + q << "])[JI" << K.obj
          << "] * " << stride << ";\n";
      else
        o << "        gid += (unsigned)(f" << K.col << " - a.iv[" << 32 + 4 * q << "]) * " << stride
          << ";\n";
    }
  }
  for (int m = 0; m < M; ++m)
    o << "        const svm_i64 mv" << m << " = (svm_i64)(" << s.metric_ctype[m] << ")(" << s.metrics[m]
      << ");\n";
}

// The aggregation of one passing row in a kernel (key, metrics, LDS / HBM accumulation).
void emit_aggregate(std::ostringstream& o, const KernelShape& s) {
  const int M = (int)s.metrics.size();
  emit_key_and_metrics(o, s);
  if (s.agg == AGG_SCALAR) {
    for (int m = 0; m < M; ++m) o << "        acc" << m << " += mv" << m << ";\n";
  } else if (s.agg == AGG_DENSE) {
    for (int m = 0; m < M; ++m)
      o << "        atomicAdd(&lsum[gid * SVM_M + " << m << "], (svm_u64)mv" << m << ");\n";
    o << "        atomicOr(&lflag[gid >> 5], 1u << (gid & 31u));\n";
  } else if (s.agg == AGG_SORT) {
    o << "        unsigned h = (unsigned)((gk * 0x9E3779B97F4A7C15ull) >> (64 - SVM_HBITS));\n"
      << "        int slot = -1;\n"
      << "        for (int t = 0; t < " << s.hash_probes << "; ++t) {\n"
      << "          const svm_u64 k = hkey64[h];\n"
      << "          if (k == gk) { slot = (int)h; break; }\n"
      << "          if (k == ~0ull) {\n"
      << "            const svm_u64 old = atomicCAS(&hkey64[h], ~0ull, gk);\n"
      << "            if (old == ~0ull || old == gk) { slot = (int)h; break; }\n"
      << "          }\n"
      << "          h = (h + 1u) & (SVM_HSLOTS - 1u);\n"
      << "        }\n"
      << "        if (slot >= 0) {\n";
    for (int m = 0; m < M; ++m)
      o << "          atomicAdd(&hsum[slot * SVM_M + " << m << "], (svm_u64)mv" << m << ");\n";
    o << "        }\n"
      << "        const svm_u64 pos = svm_wave_reserve((svm_u64*)a.out[5], slot < 0);\n"
      << "        if (slot < 0 && pos < svm_cap) {\n"
      << "          ((svm_u64*)a.out[3])[pos] = gk;\n";
    for (int m = 0; m < M; ++m)
      o << "          ((svm_i64*)a.out[4])[pos * SVM_M + " << m << "] = mv" << m << ";\n";
    o << "        }\n";
  } else {
    o << "        unsigned h = (gid * 2654435761u) >> (32 - SVM_HBITS);\n"
      << "        int slot = -1;\n"
      << "        for (int t = 0; t < " << s.hash_probes << "; ++t) {\n"
      << "          unsigned k = hkey[h];\n"
      << "          if (k == gid + 1u) { slot = (int)h; break; }\n"
      << "          if (k == 0u) {\n"
      << "            unsigned old = atomicCAS(&hkey[h], 0u, gid + 1u);\n"
      << "            if (old == 0u || old == gid + 1u) { slot = (int)h; break; }\n"
      << "          }\n"
      << "          h = (h + 1u) & (SVM_HSLOTS - 1u);\n"
      << "        }\n"
      << "        if (slot >= 0) {\n";
    for (int m = 0; m < M; ++m)
      o << "          atomicAdd(&hsum[slot * SVM_M + " << m << "], (svm_u64)mv" << m << ");\n";
    o << "        } else {\n";
    for (int m = 0; m < M; ++m)
      o << "          atomicAdd(&gsum[(size_t)gid * SVM_M + " << m << "], (svm_u64)mv" << m << ");\n";
    o << "          atomicOr(&gflag[gid >> 5], 1u << (gid & 31u));\n"
      << "        }\n";
  }
}

}  // namespace

namespace {

// One dense-map join test of the flat form for row e of group u: `cond` becomes
// false when the row's key has no slot (rows already failing read slot 0), and the
// payload (or, for a deferred map, the slot offset) is kept in pp<j>_<u>[e].
void emit_flat_join_test(std::ostringstream& o, const KernelShape& s, size_t j, int u) {
  const JoinShape& J = s.joins[j];
  if (J.defer) {
    // the existence bit decides; the slot offset is kept and the payload read
    // only for rows that pass every join
    o << "      { const unsigned off = (unsigned)(f" << J.lcol << " - a.iv[" << 4 * j << "]);\n"
      << "        const bool in = cond && off < (unsigned)a.iv[" << 4 * j + 1 << "];\n"
      << "        const unsigned b = svm_map_get<1>(mb" << j << ", in ? off : 0u);\n"
      << "        cond = in && b != 0u;\n"
      << "        pp" << j << "_" << u << "[e] = off; }\n";
    return;
  }
  o << "      { const unsigned off = (unsigned)(f" << J.lcol << " - a.iv[" << 4 * j << "]);\n"
    << "        const bool in = cond && off < (unsigned)a.iv[" << 4 * j + 1 << "];\n"
    << "        unsigned p = svm_map_get<" << J.wbits << ">(m" << j << ", in ? off : 0u);\n"
    << "        p = in ? p : 0u;\n"
    << "        cond = p != 0u;\n"
    << "        pp" << j << "_" << u << "[e] = p; }\n";
}

// The flat form's second sweep: each passing row of group u is aggregated.
void emit_flat_aggregate(std::ostringstream& o, const KernelShape& s, int u) {
  o << "    #pragma unroll\n    for (int e = 0; e < 4; ++e) {\n"
    << "      if (!cnd" << u << "[e]) continue;\n"
    << "      const long long i = v" << u << "_ * 4 + e;\n      (void)i;\n";
  for (int c : s.fact_i32) o << "      const int f" << c << " = x" << c << "_" << u << "[e];\n";
  for (size_t j = 0; j < s.joins.size(); ++j) {
    if (s.joins[j].defer)
      o << "      const unsigned p" << j << " = svm_map_get<" << s.joins[j].wbits << ">(m" << j << ", pp" << j
        << "_" << u << "[e]);\n      (void)p" << j << ";\n";
    else
      o << "      const unsigned p" << j << " = pp" << j << "_" << u << "[e];\n      (void)p" << j << ";\n";
    if (s.joins[j].payload == PAY_ROW)
      o << "      const long long JI" << s.joins[j].robj << " = (long long)p" << j << " - 1;\n";
  }
  o << "      {\n";
  emit_aggregate(o, s);
  o << "      }\n    }\n";
}

// Load tiers of the staged form: the tier of each source column and of the filter.
struct Tiers {
  bool on = false;
  int last = 0;               // the tier after every join tier (metric, key, filter columns)
  std::map<int, int> col;     // fact_i32 column -> tier
  int filter = 0;             // tier at which the filter is applied
};

// The Int32 source columns an expression names as f<c>; false when it reads anything
// else of a row (other source columns, dimension columns through a join).
bool int_columns_of(const std::string& e, std::vector<int>& cols) {
  if (e.find("FCOL") != std::string::npos || e.find("DCOL") != std::string::npos ||
      e.find("JI") != std::string::npos)
    return false;
  for (size_t i = 0; i < e.size(); ++i) {
    if (e[i] != 'f' || (i > 0 && (isalnum((unsigned char)e[i - 1]) || e[i - 1] == '_'))) continue;
    size_t k = i + 1;
    while (k < e.size() && isdigit((unsigned char)e[k])) ++k;
    if (k == i + 1 || (k < e.size() && (isalnum((unsigned char)e[k]) || e[k] == '_'))) continue;
    cols.push_back(atoi(e.c_str() + i + 1));
  }
  return true;
}

Tiers column_tiers(const KernelShape& s, bool flat) {
  Tiers t;
  if (!flat || s.join_tier.size() != s.joins.size()) return t;
  for (int jt : s.join_tier) t.last = std::max(t.last, jt + 1);
  if (t.last == 0) return t;  // no join in a tier
  t.on = true;
  for (int c : s.fact_i32) t.col[c] = t.last;
  for (size_t j = 0; j < s.joins.size(); ++j)
    if (s.join_tier[j] >= 0) t.col[s.joins[j].lcol] = std::min(t.col[s.joins[j].lcol], s.join_tier[j]);
  // the filter goes with the latest of its columns (after every tier when it reads more
  // than Int32 source columns)
  t.filter = 0;
  std::vector<int> fc;
  if (s.filter != "true") {
    if (!int_columns_of(s.filter, fc)) t.filter = t.last;
    for (int c : fc) t.filter = std::max(t.filter, t.col.count(c) ? t.col[c] : t.last);
  }
  return t;
}

// The staged form for all U groups: tier by tier (so that the groups' loads and lookups
// of a tier are in flight together) a lane loads the tier's columns (tier 0: with the
// iteration) only if one of its rows still passes, then applies the tier's joins (and
// the filter at its tier) to its four rows; then each group's passing rows are
// aggregated.  `pf_next` (the next iteration's tier-0 loads) goes after the tier-0
// lookups.
void emit_tiered(std::ostringstream& o, const KernelShape& s, const Tiers& tz, int U, bool pf,
                 const std::string& pf_next) {
  for (int u = 0; u < U; ++u) {
    o << "    bool cnd" << u << "[4];\n";
    for (size_t j = 0; j < s.joins.size(); ++j) o << "    unsigned pp" << j << "_" << u << "[4];\n";
  }
  for (int t = 0; t <= tz.last; ++t) {
    std::vector<int> load;
    for (int c : s.fact_i32)
      if (t > 0 && tz.col.at(c) == t) load.push_back(c);
    std::vector<size_t> js;
    for (size_t j = 0; j < s.joins.size(); ++j)
      if (s.join_tier[j] == t || (s.join_tier[j] < 0 && t == tz.last)) js.push_back(j);
    const bool filt = s.filter != "true" && tz.filter == t;
    const bool tests = t == 0 || !js.empty() || filt;
    auto gated_loads = [&](int u) {
      if (load.empty()) return;
      const std::string idx = pf ? std::string("v0_") : "vc" + I(u);
      o << "    if (cnd" << u << "[0] | cnd" << u << "[1] | cnd" << u << "[2] | cnd" << u << "[3]) {\n";
      for (int c : load)
        o << "      x" << c << "_" << u << " = __builtin_nontemporal_load((const svm_i32x4*)a.ptr[" << c << "] + "
          << idx << ");\n";
      o << "    }\n";
    };
    auto row_tests = [&](int u) {
      if (!tests) return;
      o << "    #pragma unroll\n    for (int e = 0; e < 4; ++e) {\n"
        << "      const long long i = v" << u << "_ * 4 + e;\n      (void)i;\n";
      for (int c : s.fact_i32)
        if (tz.col.at(c) <= t)
          o << "      const int f" << c << " = x" << c << "_" << u << "[e];\n      (void)f" << c << ";\n";
      if (t == 0)
        o << "      bool cond = (i >= r_lo) & (i < r_hi);\n";
      else
        o << "      bool cond = cnd" << u << "[e];\n";
      if (filt) o << "      cond = cond && (" << s.filter << ");\n";
      for (size_t j : js) emit_flat_join_test(o, s, j, u);
      o << "      cnd" << u << "[e] = cond;\n    }\n";
    };
    for (int u = 0; u < U; ++u) gated_loads(u);
    for (int u = 0; u < U; ++u) row_tests(u);
    if (t == 0 && !pf_next.empty()) o << "    asm volatile(\"\" ::: \"memory\");\n" << pf_next;
  }
  for (int u = 0; u < U; ++u) emit_flat_aggregate(o, s, u);
}

// The compacting staged form.  Phase 1: each lane loads U int4 groups of the tier-0
// columns and tests its rows against the tier-0 joins (and the filter when it reads
// only tier-0 columns); the passing rows are marked in a 4U-bit mask.  Phase 2: the
// wave's passing rows are numbered (prefix sum of the mask counts), lane l takes row
// number kb + l (its owner lane found by a binary search over the exclusive prefixes
// and the row by the owner's mask) and gathers the key columns of tiers 1.. one tier
// at a time while the row passes.  Phase 3: the rows that pass every tier are queued
// per wave; every 64 of them (and the rest at the end) the wave gathers their whole
// rows (filter and every join again, for the payloads) and aggregates them, one per
// lane.
void emit_compact(std::ostringstream& o, const KernelShape& s, const Tiers& tz, int U) {
  // The one tier-0 column read from its narrow copy: a lane streams groups of 8 rows
  // (a 128-bit load of their low 16 bits and, for a range beyond 16 bits, a 16-bit or
  // 64-bit load of their high bits, both coalesced over the wave's 512 rows) instead
  // of int4 groups of 4 rows.
  int c0 = -1;
  {
    std::vector<int> t0;
    for (int c : s.fact_i32)
      if (tz.col.at(c) == 0) t0.push_back(c);
    if (t0.size() == 1 && !s.packed.empty() && s.packed[0] == t0[0]) c0 = t0[0];
  }
  const int R = c0 >= 0 ? 8 : 4, RS = c0 >= 0 ? 3 : 2, HB = s.pack_high;
  if (R * U > 32) U = 32 / R;  // a lane's rows are the bits of one 32-bit mask
  // a later tier's gather of fact column c at row i
  auto gather = [&](int c) { return "((const int*)a.ptr[" + I(c) + "])[i]"; };
  o << "  const unsigned svm_lane = threadIdx.x & 63u;\n"
    // rows that passed every tier wait in a 128-entry ring per wave until 64 have
    // gathered; the wave then evaluates and aggregates them together, one per lane
    << "  __shared__ svm_u64 svm_q[SVM_BLOCK * 2];\n"
    << "  svm_u64* const svm_qw = svm_q + (threadIdx.x >> 6) * 128u;\n"
    << "  unsigned svm_qh = 0u, svm_qn = 0u;\n"
    << "  auto svm_drain = [&](unsigned cnt) {\n"
    << "    if (svm_lane < cnt) {\n"
    << "      const long long i = (long long)svm_qw[(svm_qh + svm_lane) & 127u];\n"
    << "      bool cond = true;\n";
  for (int c : s.fact_i32) o << "      const int f" << c << " = ((const int*)a.ptr[" << c << "])[i];\n";
  emit_row_body_joins(o, s, "      ");
  o << "      if (cond) {\n";
  emit_aggregate(o, s);
  o << "      }\n"
    << "    }\n"
    << "    svm_qh = (svm_qh + cnt) & 127u;\n"
    << "    svm_qn -= cnt;\n"
    << "  };\n";
  o << "  const long long g_lo = r_lo >> " << RS << ", g_hi = (r_hi + " << R - 1 << ") >> " << RS << ";\n"
    << "  const long long CH = (long long)SVM_BLOCK * " << U << ";\n"
    << "  for (long long vb = g_lo + (long long)blockIdx.x * CH; vb < g_hi; vb += (long long)gridDim.x * CH) {\n";
  for (int u = 0; u < U; ++u) {
    o << "    const long long v" << u << "_ = vb + " << u << "ll * SVM_BLOCK + threadIdx.x;\n"
      << "    const long long vc" << u << " = v" << u << "_ < g_hi ? v" << u << "_ : g_lo;\n";
    if (c0 >= 0) {
      // group g of 8 rows: block g / 64 of 1,024 + 64 x HB bytes, lane slot g % 64 of
      // each plane
      o << "    const char* svm_pb" << u << " = (const char*)a.ptr[" << SVM_SLOT_PACK << "] + " << 1024 + 64 * HB
        << "ll * (vc" << u << " >> 6);\n"
        << "    const svm_i32x4 lo" << u << " = __builtin_nontemporal_load((const svm_i32x4*)svm_pb" << u << " + (vc" << u
        << " & 63));\n";
      if (HB == 8)
        o << "    const svm_i32x2 hi" << u << " = __builtin_nontemporal_load((const svm_i32x2*)(svm_pb" << u
          << " + 1024) + (vc" << u << " & 63));\n";
      else if (HB == 2)
        o << "    const unsigned hi" << u << " = (unsigned)__builtin_nontemporal_load((const unsigned short*)(svm_pb" << u
          << " + 1024) + (vc" << u << " & 63));\n";
    } else {
      for (int c : s.fact_i32)
        if (tz.col.at(c) == 0)
          o << "    const svm_i32x4 x" << c << "_" << u << " = __builtin_nontemporal_load((const svm_i32x4*)a.ptr[" << c
            << "] + vc" << u << ");\n";
    }
  }
  o << "    unsigned svm_m = 0u;\n";
  const bool filt0 = s.filter != "true" && tz.filter == 0;
  for (int u = 0; u < U; ++u) {
    // the group's rows inside [r_lo, r_hi): all of them except in the first and last
    // groups of the range, so the rows are tested without a bound check each and the
    // group's mask is trimmed once
    o << "    unsigned svm_v" << u << ";\n"
      << "    { const long long r0 = v" << u << "_ * " << R << ";\n"
      << "      svm_v" << u << " = r0 >= r_lo && r0 + " << R << " <= r_hi ? " << (1u << R) - 1u << "u : 0u;\n"
      << "      if (svm_v" << u << " == 0u)\n"
      << "        for (int e = 0; e < " << R << "; ++e) svm_v" << u << " |= (unsigned)((r0 + e >= r_lo) & (r0 + e < r_hi)) << e; }\n";
    o << "    unsigned svm_g" << u << " = 0u;\n"
      << "    #pragma unroll\n    for (int e = 0; e < " << R << "; ++e) {\n";
    if (c0 >= 0) {
      if (HB == 8) {
        // row e of the group: halfword e of the low plane, byte e of the high plane, put
        // together by one byte permute ({high word, low word}: bytes 2e', 2e'+1 of the
        // low word, byte e'' of the high word, then a zero byte)
        o << "      const unsigned svm_sel = 0x0C000000u | ((4u + (unsigned)(e & 3)) << 16) | "
             "((2u * (unsigned)(e & 1) + 1u) << 8) | (2u * (unsigned)(e & 1));\n"
          << "      const int f" << c0 << " = (int)__builtin_amdgcn_perm((unsigned)hi" << u << "[e >> 2], (unsigned)lo" << u
          << "[e >> 1], svm_sel) + a.iv[48];\n      (void)f" << c0 << ";\n";
      } else {
        // row e of the group: halfword e of the low plane (a byte permute with zero
        // bytes above it) and, for HB = 2, bits 2e, 2e + 1 of the high word as bits 16, 17
        o << "      const unsigned svm_sel = 0x0C0C0000u | ((2u * (unsigned)(e & 1) + 1u) << 8) | (2u * (unsigned)(e & 1));\n"
          << "      const int f" << c0 << " = (int)(__builtin_amdgcn_perm(0u, (unsigned)lo" << u << "[e >> 1], svm_sel)";
        if (HB == 2) o << " | (((hi" << u << " >> (2 * e)) & 3u) << 16)";
        o << ") + a.iv[48];\n      (void)f" << c0 << ";\n";
      }
    } else {
      for (int c : s.fact_i32)
        if (tz.col.at(c) == 0) o << "      const int f" << c << " = x" << c << "_" << u << "[e];\n      (void)f" << c << ";\n";
    }
    // every row's tests run without branches: a key outside the map's range is clamped
    // to slot `range`, whose existence bit is 0; a map lookup reads slot 0 instead
    o << "      unsigned svm_b = 1u;\n";
    if (filt0) o << "      svm_b &= (" << s.filter << ") ? 1u : 0u;\n";
    for (size_t j = 0; j < s.joins.size(); ++j) {
      if (s.join_tier[j] != 0) continue;
      const JoinShape& J = s.joins[j];
      if (J.defer) {
        o << "      { const unsigned off = min((unsigned)(f" << J.lcol << " - a.iv[" << 4 * j << "]), (unsigned)a.iv["
          << 4 * j + 1 << "]);\n"
          << "        svm_b &= svm_map_get<1>(mb" << j << ", off); }\n";
      } else {
        o << "      { const unsigned off = (unsigned)(f" << J.lcol << " - a.iv[" << 4 * j << "]);\n"
          << "        const bool in = off < (unsigned)a.iv[" << 4 * j + 1 << "];\n"
          << "        const unsigned p = svm_map_get<" << J.wbits << ">(m" << j << ", in ? off : 0u);\n"
          << "        svm_b &= (in && p != 0u) ? 1u : 0u; }\n";
      }
    }
    o << "      svm_g" << u << " |= svm_b << e;\n    }\n"
      << "    svm_m |= (svm_g" << u << " & svm_v" << u << ") << " << R * u << ";\n";
  }
  o << "    const unsigned svm_c = (unsigned)__popc(svm_m);\n"
    << "    unsigned svm_in = svm_c;\n"
    << "    for (int d = 1; d < 64; d <<= 1) {\n"
    << "      const unsigned t = __shfl_up(svm_in, d, 64);\n"
    << "      if (svm_lane >= (unsigned)d) svm_in += t;\n"
    << "    }\n"
    << "    const unsigned svm_ex = svm_in - svm_c;\n"
    << "    const unsigned svm_tot = __shfl(svm_in, 63, 64);\n"
    << "    for (unsigned kb = 0; kb < svm_tot; kb += 64u) {\n"
    << "      const unsigned k = kb + svm_lane;\n"
    << "      int L = 0;\n"
    << "      for (int st = 32; st > 0; st >>= 1) {\n"
    << "        const unsigned ex = __shfl(svm_ex, L + st, 64);\n"
    << "        if (ex <= k) L += st;\n"
    << "      }\n"
    << "      const unsigned mo = __shfl(svm_m, L, 64);\n"
    << "      const unsigned nth = k - __shfl(svm_ex, L, 64);\n"
    << "      bool fin = false;\n"
    << "      long long fi = 0;\n"
    << "      if (k < svm_tot) {\n"
    << "        unsigned mm = mo;\n"
    << "        for (unsigned t = 0; t < nth; ++t) mm &= mm - 1u;\n"
    << "        const unsigned bit = (unsigned)__builtin_ctz(mm);\n"
    << "        const long long i = (vb + (long long)(bit >> " << RS << ") * SVM_BLOCK + (long long)((threadIdx.x & ~63u) + "
       "(unsigned)L)) * " << R << " + (long long)(bit & " << R - 1 << "u);\n"
    << "        bool cond = true;\n";
  for (int t = 1; t < tz.last; ++t) {
    o << "        if (cond) {\n";
    for (size_t j = 0; j < s.joins.size(); ++j) {
      if (s.join_tier[j] != t) continue;
      const JoinShape& J = s.joins[j];
      o << "          { const unsigned off = (unsigned)(" << gather(J.lcol) << " - a.iv[" << 4 * j << "]);\n";
      if (J.defer)
        o << "            cond = cond && off < (unsigned)a.iv[" << 4 * j + 1 << "] && svm_map_get<1>(mb" << j
          << ", off) != 0u; }\n";
      else
        o << "            cond = cond && off < (unsigned)a.iv[" << 4 * j + 1 << "] && svm_map_get<" << J.wbits << ">(m"
          << j << ", off) != 0u; }\n";
    }
    o << "        }\n";
  }
  o << "        fin = cond;\n"
    << "        fi = i;\n"
    << "      }\n"
    << "      const unsigned long long fb = __ballot(fin);\n"
    << "      if (fb) {\n"
    << "        const unsigned p = svm_qn + (unsigned)__popcll(fb & ((1ull << svm_lane) - 1ull));\n"
    << "        if (fin) svm_qw[(svm_qh + p) & 127u] = (svm_u64)fi;\n"
    << "        svm_qn += (unsigned)__popcll(fb);\n"
    << "        __builtin_amdgcn_wave_barrier();\n"
    << "        if (svm_qn >= 64u) svm_drain(64u);\n"
    << "      }\n"
    << "    }\n"
    << "  }\n"
    << "  while (svm_qn > 0u) svm_drain(svm_qn < 64u ? svm_qn : 64u);\n";
}

}  // namespace

std::vector<int> compact_packable(const KernelShape& s) {
  std::vector<int> out;
  bool flat = s.flat;
  for (auto& J : s.joins) flat = flat && J.kind == JOIN_MAP;
  const Tiers tz = column_tiers(s, flat);
  if (!tz.on || !s.compact) return out;
  std::vector<int> t0;
  for (int c : s.fact_i32)
    if (tz.col.at(c) == 0) t0.push_back(c);
  if (t0.size() != 1 || (s.filter != "true" && tz.filter == 0)) return out;
  out.push_back(t0[0]);
  return out;
}

std::string gen_scan_source(const KernelShape& s, const char* entry) {
  std::ostringstream o;
  emit_prologue(o, s);
  const int M = (int)s.metrics.size();
  const int U = s.unroll < 1 ? 1 : s.unroll;
  o << "#define SVM_M " << M << "\n";
  if (s.agg == AGG_HASH || s.agg == AGG_SORT) {
    o << "#define SVM_HSLOTS " << (1 << s.hash_bits) << "\n";
    o << "#define SVM_HBITS " << s.hash_bits << "\n";
  }
  o << "extern \"C\" __global__ void __launch_bounds__(SVM_BLOCK) " << entry << "(SvmArgs a) {\n";
  o << "  extern __shared__ svm_u64 svm_lds64[];\n  unsigned* svm_lds = (unsigned*)svm_lds64;\n";
  emit_colptrs(o, s);
  emit_stage_maps(o, s);
  emit_map_ptrs(o, s);
  o << "  const unsigned G = (unsigned)a.iv[44];\n  (void)G;\n";
  if (s.agg == AGG_DENSE) {
    o << "  svm_u64* lsum = (svm_u64*)(svm_lds + a.iv[45]);\n"
      << "  unsigned* lflag = (unsigned*)(lsum + (size_t)G * SVM_M);\n"
      << "  for (unsigned g = threadIdx.x; g < G * SVM_M; g += SVM_BLOCK) lsum[g] = 0ull;\n"
      << "  for (unsigned w = threadIdx.x; w < (G + 31u) / 32u; w += SVM_BLOCK) lflag[w] = 0u;\n";
  } else if (s.agg == AGG_SORT) {
    o << "  svm_u64* hkey64 = (svm_u64*)(svm_lds + a.iv[45]);\n"
      << "  svm_u64* hsum = hkey64 + SVM_HSLOTS;\n"
      << "  for (unsigned q = threadIdx.x; q < SVM_HSLOTS; q += SVM_BLOCK) hkey64[q] = ~0ull;\n"
      << "  for (unsigned q = threadIdx.x; q < SVM_HSLOTS * SVM_M; q += SVM_BLOCK) hsum[q] = 0ull;\n"
      << "  const svm_u64 svm_cap = (svm_u64)(unsigned)a.iv[46] | ((svm_u64)(unsigned)a.iv[47] << 32);\n";
  } else if (s.agg == AGG_HASH) {
    o << "  unsigned* hkey = svm_lds + a.iv[45];\n"
      << "  svm_u64* hsum = (svm_u64*)(hkey + SVM_HSLOTS);\n"
      << "  for (unsigned q = threadIdx.x; q < SVM_HSLOTS; q += SVM_BLOCK) hkey[q] = 0u;\n"
      << "  for (unsigned q = threadIdx.x; q < SVM_HSLOTS * SVM_M; q += SVM_BLOCK) hsum[q] = 0ull;\n";
  }
  o << "  svm_u64* gsum = (svm_u64*)a.out[0];\n  unsigned* gflag = (unsigned*)a.out[1];\n"
    << "  (void)gsum; (void)gflag;\n";
  o << "  __syncthreads();\n";
  if (s.agg == AGG_SCALAR)
    for (int m = 0; m < M; ++m) o << "  svm_i64 acc" << m << " = 0;\n";
  o << "  const long long r_lo = a.row0, r_hi = a.row0 + a.rows;\n"
    << "  const long long v_lo = r_lo >> 2, v_hi = (r_hi + 3) >> 2;\n"
    << "  const long long GT = (long long)gridDim.x * SVM_BLOCK;\n";
  const bool pf = s.prefetch && U == 1;
  // the flat form holds dense-map lookups only; probe joins use the per-row form
  bool flat = s.flat;
  for (auto& J : s.joins) flat = flat && J.kind == JOIN_MAP;
  const Tiers tz = column_tiers(s, flat);
  const bool compact = tz.on && s.compact;
  if (compact) emit_compact(o, s, tz, U > 8 ? 8 : U);
  if (!compact) {
    // columns loaded at the top of an iteration (all of them without staged loads)
    auto up_front = [&](int c) { return !tz.on || tz.col.at(c) == 0; };
    std::string pf_next;
    if (pf) {
      // software pipelining: the next iteration's 128-bit groups are requested before
      // this iteration's rows are probed and aggregated (out-of-range indexes are
      // clamped to v_lo, a valid group whose rows fail the range test)
      o << "  long long v0 = v_lo + (long long)blockIdx.x * SVM_BLOCK + threadIdx.x;\n";
      for (int c : s.fact_i32)
        if (up_front(c))
          o << "  svm_i32x4 n" << c << " = __builtin_nontemporal_load((const svm_i32x4*)a.ptr[" << c
            << "] + (v0 < v_hi ? v0 : v_lo));\n";
      o << "  for (; v0 < v_hi; v0 += GT) {\n"
        << "    const long long v0_ = v0;\n";
      for (int c : s.fact_i32) {
        if (up_front(c))
          o << "    const svm_i32x4 x" << c << "_0 = n" << c << ";\n";
        else
          o << "    svm_i32x4 x" << c << "_0 = {0, 0, 0, 0};\n";
      }
      std::ostringstream nx;
      nx << "    { const long long vn = v0 + GT; const long long vnc = vn < v_hi ? vn : v_lo;\n";
      for (int c : s.fact_i32)
        if (up_front(c))
          nx << "      n" << c << " = __builtin_nontemporal_load((const svm_i32x4*)a.ptr[" << c << "] + vnc);\n";
      nx << "    }\n";
      // staged: requested after this iteration's first map lookups (loads return in
      // order, so a lookup issued after the prefetch would wait for it)
      if (tz.on)
        pf_next = nx.str();
      else
        o << nx.str();
    } else {
      o << "  for (long long v0 = v_lo + (long long)blockIdx.x * SVM_BLOCK + threadIdx.x; v0 < v_hi; v0 += "
        << U << " * GT) {\n";
      for (int u = 0; u < U; ++u) {
        o << "    const long long v" << u << "_ = v0 + " << u << " * GT;\n"
          << "    const long long vc" << u << " = v" << u << "_ < v_hi ? v" << u << "_ : v_lo;\n";
        for (int c : s.fact_i32) {
          if (up_front(c))
            o << "    const svm_i32x4 x" << c << "_" << u
              << " = __builtin_nontemporal_load((const svm_i32x4*)a.ptr[" << c << "] + vc" << u << ");\n";
          else
            o << "    svm_i32x4 x" << c << "_" << u << " = {0, 0, 0, 0};\n";
        }
      }
    }
    if (tz.on) emit_tiered(o, s, tz, U, pf, pf_next);
    for (int u = 0; u < U && !tz.on; ++u) {
      if (!flat) {
      o << "    #pragma unroll\n    for (int e = 0; e < 4; ++e) {\n"
        << "      const long long i = v" << u << "_ * 4 + e;\n";
      for (int c : s.fact_i32) o << "      const int f" << c << " = x" << c << "_" << u << "[e];\n";
      o << "      bool cond = (i >= r_lo) & (i < r_hi);\n";
      emit_row_body_joins(o, s, "      ");
      o << "      if (cond) {\n";
      emit_aggregate(o, s);
      o << "      }\n    }\n";
      } else {
      // flat: the four rows' filters and map lookups are straight-line code (a row
      // that needs no lookup reads entry 0), so their loads are in flight together;
      // the passing rows are then aggregated one by one
      o << "    bool cnd" << u << "[4];\n";
      for (size_t j = 0; j < s.joins.size(); ++j) o << "    unsigned pp" << j << "_" << u << "[4];\n";
      o << "    #pragma unroll\n    for (int e = 0; e < 4; ++e) {\n"
        << "      const long long i = v" << u << "_ * 4 + e;\n";
      for (int c : s.fact_i32) o << "      const int f" << c << " = x" << c << "_" << u << "[e];\n";
      o << "      bool cond = (i >= r_lo) & (i < r_hi);\n"
        << "      cond = cond && (" << s.filter << ");\n";
      for (size_t j = 0; j < s.joins.size(); ++j) emit_flat_join_test(o, s, j, u);
      o << "      cnd" << u << "[e] = cond;\n    }\n";
      emit_flat_aggregate(o, s, u);
      }
    }
    o << "  }\n";
  }
  // flush
  if (s.agg == AGG_SCALAR) {
    o << "  __shared__ svm_i64 red[SVM_BLOCK / 64][" << (M > 0 ? M : 1) << "];\n";
    for (int m = 0; m < M; ++m) {
      o << "  for (int d = 32; d > 0; d >>= 1) acc" << m << " += __shfl_xor(acc" << m << ", d);\n"
        << "  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][" << m << "] = acc" << m << ";\n";
    }
    o << "  __syncthreads();\n"
      << "  if (threadIdx.x < " << M << ") {\n"
      << "    svm_i64 t = 0;\n"
      << "    for (int w = 0; w < SVM_BLOCK / 64; ++w) t += red[w][threadIdx.x];\n"
      << "    atomicAdd((svm_u64*)a.out[2] + threadIdx.x, (svm_u64)t);\n"
      << "  }\n";
  } else if (s.agg == AGG_DENSE) {
    o << "  __syncthreads();\n"
      << "  for (unsigned g = threadIdx.x; g < G; g += SVM_BLOCK)\n"
      << "    if ((lflag[g >> 5] >> (g & 31u)) & 1u) {\n";
    for (int m = 0; m < M; ++m)
      o << "      atomicAdd(&gsum[(size_t)g * SVM_M + " << m << "], lsum[g * SVM_M + " << m << "]);\n";
    o << "    }\n"
      << "  for (unsigned w = threadIdx.x; w < (G + 31u) / 32u; w += SVM_BLOCK)\n"
      << "    if (lflag[w]) atomicOr(&gflag[w], lflag[w]);\n";
  } else if (s.agg == AGG_SORT) {
    // every occupied LDS slot becomes one appended entry (SVM_HSLOTS is a multiple of
    // SVM_BLOCK, so every lane runs the same iterations)
    o << "  __syncthreads();\n"
      << "  for (unsigned q = threadIdx.x; q < SVM_HSLOTS; q += SVM_BLOCK) {\n"
      << "    const svm_u64 k = hkey64[q];\n"
      << "    const svm_u64 pos = svm_wave_reserve((svm_u64*)a.out[5], k != ~0ull);\n"
      << "    if (k != ~0ull && pos < svm_cap) {\n"
      << "      ((svm_u64*)a.out[3])[pos] = k;\n";
    for (int m = 0; m < M; ++m)
      o << "      ((svm_i64*)a.out[4])[pos * SVM_M + " << m << "] = (svm_i64)hsum[q * SVM_M + " << m << "];\n";
    o << "    }\n  }\n";
  } else {
    o << "  __syncthreads();\n"
      << "  for (unsigned q = threadIdx.x; q < SVM_HSLOTS; q += SVM_BLOCK) {\n"
      << "    const unsigned k = hkey[q];\n"
      << "    if (k) {\n"
      << "      const unsigned g = k - 1u;\n";
    for (int m = 0; m < M; ++m)
      o << "      atomicAdd(&gsum[(size_t)g * SVM_M + " << m << "], hsum[q * SVM_M + " << m << "]);\n";
    o << "      atomicOr(&gflag[g >> 5], 1u << (g & 31u));\n"
      << "    }\n  }\n";
  }
  o << "}\n";
  return o.str();
}

// ------------------------------------------------------------------ source memo
// Generating a kernel's source text takes ~20 us; a pipeline shape's text is the
// same every run (constants travel as launch parameters), so it is memoised under a
// compact description of the shape, with its hash.
namespace {

// Every field of the shape, strings length-prefixed, so that two shapes share a key
// only when they are equal.
void put(std::ostringstream& k, const std::string& v) { k << v.size() << ':' << v; }

std::string shape_key(const KernelShape& s, char kind, const char* entry) {
  static_assert(sizeof(KernelShape) == 368 && sizeof(JoinShape) == 36 && sizeof(KeyShape) == 16,
                "a field was added to the kernel shape: add it to shape_key");
  std::ostringstream k;
  k << kind << entry << '|' << s.block << ',' << s.unroll << ',' << s.prefetch << s.flat << ',' << s.agg << ','
    << s.hash_bits << ',' << s.hash_probes << '|';
  put(k, s.filter);
  for (int c : s.fact_i32) k << c << ',';
  k << '|';
  for (int c : s.fact_other) k << c << ',';
  k << '|';
  for (auto& J : s.joins)
    k << J.robj << ',' << J.lcol << ',' << J.payload << ',' << J.wbits << ',' << J.lds << J.defer << J.bits_lds << ',' << J.kind << ','
      << J.rcol << ',' << J.ktype << ',' << J.ksize << ';';
  k << '|';
  for (auto& kv : s.dcol_slot) k << kv.first.first << '.' << kv.first.second << '=' << kv.second << ',';
  k << '|';
  for (auto& K : s.keys) k << K.src << ',' << K.join << ',' << K.obj << ',' << K.col << ';';
  k << '|' << s.metrics.size() << ':';
  for (auto& m : s.metrics) put(k, m);
  k << '|' << s.metric_ctype.size() << ':';
  for (auto& m : s.metric_ctype) put(k, m);
  k << '|' << s.proj.size() << ':';
  for (auto& q : s.proj) put(k, q);
  k << '|';
  for (auto& t : s.proj_type) k << t.first << ',' << t.second << ';';
  k << '|';
  for (int t : s.join_tier) k << t << ',';
  k << '|' << s.compact << '|';
  for (int c : s.packed) k << c << ',';
  k << '|' << s.pack_high;
  return k.str();
}

std::mutex g_memo_mu;
std::map<std::string, std::shared_ptr<const GeneratedSource>> g_memo;

template <typename Gen>
std::shared_ptr<const GeneratedSource> memo(const std::string& key, Gen&& gen) {
  {
    std::lock_guard<std::mutex> g(g_memo_mu);
    auto it = g_memo.find(key);
    if (it != g_memo.end()) return it->second;
  }
  auto src = std::make_shared<GeneratedSource>();
  src->text = gen();
  src->hash = source_hash(src->text);
  std::lock_guard<std::mutex> g(g_memo_mu);
  if (g_memo.size() > 4096) g_memo.clear();  // bounded: shapes are few in practice
  return g_memo.emplace(key, src).first->second;
}

}  // namespace

std::shared_ptr<const GeneratedSource> scan_source(const KernelShape& s, const char* entry) {
  return memo(shape_key(s, 'r', entry), [&] { return gen_scan_source(s, entry); });
}

std::shared_ptr<const GeneratedSource> select_source(const KernelShape& s) {
  return memo(shape_key(s, 's', ""), [&] { return gen_select_source(s); });
}

std::string gen_host_source(const KernelShape& s0) {
  // The host engine's row loop over rows [r_lo, r_hi) of the source: the same filter,
  // join and key/metric text as the kernels (maps read from host copies), each
  // passing row accumulated into the task's SvmHostOut (or, for a select, projected
  // into the task's output columns in row order), as HostQueryTransform's loop does
  // (codegen/template.cpp:61-73).
  KernelShape s = s0;
  for (auto& J : s.joins) J.lds = J.bits_lds = false;
  const int M = (int)s.metrics.size();
  const bool select = !s.proj.empty();
  std::ostringstream o;
  o << SVM_KARGS_TEXT << "\n" << SVM_HOSTOUT_TEXT << "\n" << host_runtime_text() << "\n";
  o << "#define SVM_M " << (M > 0 ? M : 1) << "\n";
  o << "extern \"C\" void svm_host_task(const SvmArgs* ap, long long r_lo, long long r_hi, SvmHostOut* out) {\n"
    << "  const SvmArgs& a = *ap;\n";
  emit_colptrs(o, s);
  emit_map_ptrs(o, s);
  o << "  const unsigned G = (unsigned)a.iv[44];\n  (void)G;\n";
  o << "  for (long long i = r_lo; i < r_hi; ++i) {\n";
  for (int c : s.fact_i32) o << "    const int f" << c << " = ((const int*)a.ptr[" << c << "])[i];\n";
  o << "    bool cond = true;\n";
  emit_row_body_joins(o, s, "    ");
  o << "    if (!cond) continue;\n";
  if (select) {
    o << "    const unsigned long long at = out->nsel++;\n";
    for (size_t q = 0; q < s.proj.size(); ++q) {
      const int t = s.proj_type[q].first;
      const long long size = s.proj_type[q].second;
      const std::string dst = "(out->sel[" + I((int)q) + "] + at * " + std::to_string(size) + "ull)";
      if (t == 7) {
        o << "    StrCopy(" << dst << ", (const char*)(" << s.proj[q] << "), " << size << ");\n";
      } else {
        const char* ct = t == 0 ? "char" : t == 1 ? "short" : t == 3 ? "long long" : "int";
        o << "    *(" << ct << "*)" << dst << " = (" << ct << ")(" << s.proj[q] << ");\n";
      }
    }
  } else {
    o << "    {\n";
    emit_key_and_metrics(o, s);
    if (s.agg == AGG_SCALAR) {
      for (int m = 0; m < M; ++m) o << "        out->acc[" << m << "] += mv" << m << ";\n";
    } else if (s.agg == AGG_SORT) {
      o << "        const unsigned long long e = out->nentries++;\n"
        << "        out->ekeys[e] = gk;\n";
      for (int m = 0; m < M; ++m) o << "        out->evals[e * SVM_M + " << m << "] = mv" << m << ";\n";
    } else {
      o << "        if (out->shared) {\n";
      for (int m = 0; m < M; ++m)
        o << "          __atomic_fetch_add(&out->gsum[(svm_u64)gid * SVM_M + " << m << "], (svm_u64)mv" << m
          << ", __ATOMIC_RELAXED);\n";
      o << "          __atomic_fetch_or(&out->gflag[gid >> 5], 1u << (gid & 31u), __ATOMIC_RELAXED);\n"
        << "        } else {\n";
      for (int m = 0; m < M; ++m)
        o << "          out->gsum[(svm_u64)gid * SVM_M + " << m << "] += (svm_u64)mv" << m << ";\n";
      o << "          out->gflag[gid >> 5] |= 1u << (gid & 31u);\n"
        << "        }\n";
    }
    o << "    }\n";
  }
  o << "  }\n}\n";
  return o.str();
}

std::string gen_select_source(const KernelShape& s) {
  std::ostringstream o;
  emit_prologue(o, s);
  o << "#define SVM_RPT 4\n";
  // A workgroup compacts one tile of SVM_BLOCK * SVM_RPT rows in row order:
  //   svm_sel_count / svm_sel_write   two launches around a scan of the tile counts;
  //   svm_sel_single                  one launch: tiles are taken from a ticket and
  //                                   each finds the rows of the earlier tiles by a
  //                                   decoupled look-back over per-tile status words.
  const char* names[3] = {"svm_sel_count", "svm_sel_write", "svm_sel_single"};
  for (int pass = 0; pass < 3; ++pass) {
    o << "extern \"C\" __global__ void __launch_bounds__(SVM_BLOCK) " << names[pass] << "(SvmArgs a) {\n";
    o << "  extern __shared__ svm_u64 svm_lds64[];\n  unsigned* svm_lds = (unsigned*)svm_lds64;\n";
    emit_colptrs(o, s);
    emit_stage_maps(o, s);
    emit_map_ptrs(o, s);
    o << "  __shared__ unsigned wcnt[SVM_RPT][SVM_BLOCK / 64];\n";
    if (pass == 2) {
      // status word: epoch << 32 | inclusive << 31 | rows (< 2^31); a word of another
      // epoch is not yet published.  The ticket counts the tiles taken; the workgroup
      // that takes the last one puts it back to 0 for the next launch on the stream.
      o << "  __shared__ unsigned long long s_tile, s_pos;\n"
        << "  __shared__ int s_err;\n"
        << "  const unsigned long long epoch = (unsigned long long)(unsigned)a.iv[46];\n"
        << "  unsigned long long* ticket = (unsigned long long*)a.out[7];\n"
        << "  unsigned long long* status = ticket + 1;\n"
        << "  const unsigned long long ntiles = ((unsigned long long)a.rows + SVM_BLOCK * SVM_RPT - 1) / (SVM_BLOCK * SVM_RPT);\n"
        << "  if (threadIdx.x == 0) {\n"
        << "    const unsigned long long t = __hip_atomic_fetch_add(ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
        << "    if (t + 1 == ntiles) __hip_atomic_store(ticket, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
        << "    s_tile = t;\n"
        << "    s_err = t >= ntiles;\n"
        << "    if (s_err) *(volatile unsigned long long*)((unsigned long long*)a.out[5] + 3) = 1ull;\n"
        << "  }\n"
        << "  __syncthreads();\n"
        << "  if (s_err) return;\n"
        << "  const unsigned long long tile = s_tile;\n";
    } else {
      o << "  __syncthreads();\n"
        << "  const unsigned long long tile = blockIdx.x;\n";
    }
    o << "  const long long r_hi = a.row0 + a.rows;\n"
      << "  const long long t0 = a.row0 + (long long)tile * (SVM_BLOCK * SVM_RPT);\n"
      << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n"
      << "  unsigned passbits = 0u;\n";
    // first sweep: conditions and per-wave counts.  A row past the end reads the last
    // row instead (t0 < r_hi for every tile), so every row's loads are unconditional and
    // the SVM_RPT rows' loads are in flight together; its condition is then cleared.
    o << "  #pragma unroll\n"
      << "  for (int k = 0; k < SVM_RPT; ++k) {\n"
      << "    const long long iu = t0 + (long long)k * SVM_BLOCK + threadIdx.x;\n"
      << "    const bool inr = iu < r_hi;\n"
      << "    const long long i = inr ? iu : r_hi - 1;\n"
      << "    bool cond = true;\n";
    for (int c : s.fact_i32)
      o << "    const int f" << c << " = ((const int*)a.ptr[" << c << "])[i];\n";
    emit_row_body_joins(o, s, "    ");
    o << "    cond = cond && inr;\n"
      << "    const unsigned long long mask = __ballot(cond);\n"
      << "    if (lane == 0) wcnt[k][wave] = (unsigned)__popcll(mask);\n"
      << "    passbits |= (cond ? 1u : 0u) << k;\n"
      << "  }\n"
      << "  __syncthreads();\n";
    if (pass == 0) {
      o << "  if (threadIdx.x == 0) {\n"
        << "    unsigned t = 0;\n"
        << "    for (int k = 0; k < SVM_RPT; ++k) for (int w = 0; w < SVM_BLOCK / 64; ++w) t += wcnt[k][w];\n"
        << "    ((unsigned*)a.out[7])[blockIdx.x] = t;\n"
        << "  }\n"
        << "}\n";
      continue;
    }
    if (pass == 1) {
      o << "  unsigned long long pos = ((const unsigned long long*)a.out[6])[blockIdx.x];\n";
    } else {
      // publish this tile's count, then wave 0 adds the counts of earlier tiles, 64
      // statuses per step, up to the nearest one that holds its inclusive prefix
      // (positions before tile 0 read as an inclusive 0)
      o << "  if (wave == 0) {\n"
        << "    unsigned long long c = 0;\n"
        << "    for (int k = 0; k < SVM_RPT; ++k) for (int w = 0; w < SVM_BLOCK / 64; ++w) c += wcnt[k][w];\n"
        << "    if (lane == 0 && tile > 0)\n"
        << "      __hip_atomic_store(status + tile, (epoch << 32) | c, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);\n"
        << "    unsigned long long pre = 0;\n"
        << "    int err = 0;\n"
        << "    if (tile > 0) {\n"
        << "      long long top = (long long)tile - 1;\n"
        << "      unsigned spins = 0;\n"
        << "      for (;;) {\n"
        << "        const long long p = top - lane;\n"
        << "        const unsigned long long v = p >= 0 ? __hip_atomic_load(status + p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)\n"
        << "                                            : ((epoch << 32) | 0x80000000ull);\n"
        << "        const bool ready = (v >> 32) == epoch;\n"
        << "        const unsigned long long rm = __ballot(ready), im = __ballot(ready && (v & 0x80000000ull));\n"
        << "        // lanes 0..k up to the nearest inclusive status k (all 64 when there is none)\n"
        << "        const unsigned long long need = im ? ((im & (~im + 1ull)) << 1) - 1ull : ~0ull;\n"
        << "        if ((rm & need) == need) {\n"
        << "          unsigned long long x = ((need >> lane) & 1ull) ? (v & 0x7FFFFFFFull) : 0ull;\n"
        << "          for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d);\n"
        << "          pre += x;\n"
        << "          if (im) break;\n"
        << "          top -= 64;\n"
        << "          continue;\n"
        << "        }\n"
        << "        // bounded: an earlier tile that never publishes ends the launch with an error\n"
        << "        if (++spins > (1u << 22)) { err = 1; break; }\n"
        << "        __builtin_amdgcn_s_sleep(1);\n"
        << "      }\n"
        << "    }\n"
        << "    if (lane == 0) {\n"
        << "      if (err) {\n"
        << "        *(volatile unsigned long long*)((unsigned long long*)a.out[5] + 3) = 1ull;\n"
        << "        __hip_atomic_store(status + tile, (epoch << 32) | 0x80000000ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);\n"
        << "      } else {\n"
        << "        __hip_atomic_store(status + tile, (epoch << 32) | 0x80000000ull | (pre + c), __ATOMIC_RELEASE,\n"
        << "                           __HIP_MEMORY_SCOPE_AGENT);\n"
        << "        if (tile + 1 == ntiles) *(volatile unsigned long long*)a.out[5] = pre + c;\n"
        << "      }\n"
        << "      s_err = err;\n"
        << "      s_pos = pre;\n"
        << "    }\n"
        << "  }\n"
        << "  __syncthreads();\n"
        << "  if (s_err) return;  // no rows are written from an inconsistent prefix\n"
        << "  unsigned long long pos = s_pos;\n";
    }
    o << "  #pragma unroll\n"
      << "  for (int k = 0; k < SVM_RPT; ++k) {\n"
      << "    unsigned before = 0;\n"
      << "    for (int w = 0; w < wave; ++w) before += wcnt[k][w];\n"
      << "    unsigned round = 0;\n"
      << "    for (int w = 0; w < SVM_BLOCK / 64; ++w) round += wcnt[k][w];\n"
      << "    const bool cond = (passbits >> k) & 1u;\n"
      << "    const unsigned long long mask = __ballot(cond);\n"
      << "    if (cond) {\n"
      << "      const long long i = t0 + (long long)k * SVM_BLOCK + threadIdx.x;\n"
      << "      const unsigned long long at = pos + before + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));\n";
    for (int c : s.fact_i32) o << "      const int f" << c << " = ((const int*)a.ptr[" << c << "])[i];\n";
    // recompute join payloads for the projection (row ids)
    for (size_t j = 0; j < s.joins.size(); ++j) {
      const JoinShape& J = s.joins[j];
      if (J.payload != PAY_ROW) continue;
      if (J.kind != JOIN_MAP) {
        o << "      " << probe_stmt(J, j, "rj" + I((int)j)) << "\n"
          << "      const long long JI" << J.robj << " = rj" << j << ";\n";
        continue;
      }
      o << "      const unsigned offj" << j << " = (unsigned)(f" << J.lcol << " - a.iv[" << 4 * j << "]);\n"
        << "      const long long JI" << J.robj << " = (long long)svm_map_get<" << J.wbits << ">(m" << j
        << ", offj" << j << ") - 1;\n";
    }
    for (size_t q = 0; q < s.proj.size(); ++q) {
      int t = s.proj_type[q].first;
      long long size = s.proj_type[q].second;
      std::string dst = "((char*)a.ptr[" + I(SVM_SLOT_OUT + (int)q) + "])";
      if (t == 7) {
        o << "      svm_str_copy(" << dst << " + at * " << size << "ull, (const char*)(" << s.proj[q] << "), "
          << size << ");\n";
      } else {
        const char* ct = t == 0 ? "char" : t == 1 ? "short" : t == 3 ? "long long" : "int";
        o << "      *(" << ct << "*)(" << dst << " + at * " << size << "ull) = (" << ct << ")(" << s.proj[q]
          << ");\n";
      }
    }
    o << "    }\n"
      << "    pos += round;\n"
      << "  }\n";
    // the last tile ends the output: the column padding behind the last row (pos is
    // then the output's row count; an output sized for its rows has it zeroed already)
    o << "  if (tile + 1 == " << (pass == 2 ? "ntiles" : "(unsigned long long)gridDim.x")
      << " && threadIdx.x < 256) {\n";
    for (size_t q = 0; q < s.proj.size(); ++q)
      o << "    ((char*)a.ptr[" << SVM_SLOT_OUT + (int)q << "])[pos * " << s.proj_type[q].second
        << "ull + threadIdx.x] = 0;\n";
    o << "  }\n";
    o << "}\n";
  }
  return o.str();
}

}  // namespace svm
