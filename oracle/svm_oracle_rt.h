/*
 * svm_oracle_rt.h — CPU oracle runtime.  TEST INFRASTRUCTURE ONLY: nothing in the
 * product (shadowvm_amd/) may include, link or call this; only tests/, the smoke
 * check and bench.py's cpu_baseline leg use it, as the checker.
 *
 * A restatement of the reference's *generated CPU path*:
 *   - the runtime helpers included into generated host code
 *       Hash / StrHash / StrCmp / StrCopy        com/cpp/cpu-runtime.hpp:37-78
 *       InsertHashTable<T> / <char>              com/cpp/cpu-runtime.hpp:80-103
 *       Probe<T> / Probe<char>                   com/cpp/cpu-runtime.hpp:125-152
 *       BuildHashTable<T>                        com/cpp/com.hpp:168-174
 *   - the hash-table build of Pipeline::Open     core/pipeline.cu:82-137
 *   - one task per bulk (Pipeline::Run / TransformTask) core/pipeline.cu:139-344
 *   - per-bulk sort + reduce_by_key (ReduceTask) core/pipeline.cu:360-472
 * with the documented deviations: D1 metric sums accumulate in int64; D2 per-bulk
 * partial aggregates are merged into one key-ordered result (select outputs are
 * concatenated in bulk order); the oracle counts inserts that the 10-probe cap drops.
 *
 * The per-row loop itself is generated per plan by oracle/oracle.py (restating
 * codegen/codegen.cpp:100-277 and codegen/template.cpp:61-73) and includes this file.
 */
#ifndef SVM_ORACLE_RT_H_
#define SVM_ORACLE_RT_H_

#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

typedef unsigned long long uint64;

static const uint64 HashRate = 5; /* com/cpp/define.hpp:29 */

/* Hash31 over n bytes read as (signed) char, uint32 wrap, top bit cleared.
 * com/cpp/cpu-runtime.hpp:68-74 */
inline unsigned Hash(const char* p, int n) {
  unsigned h = 0;
  for (int k = 0; k < n; ++k) h = h * 31u + (unsigned)(int)(signed char)p[k];
  return h & 0x7FFFFFFFu;
}

/* As Hash but stops at the first NUL.  com/cpp/cpu-runtime.hpp:60-66 */
inline unsigned StrHash(const char* p, int n) {
  unsigned h = 0;
  for (int k = 0; k < n && p[k] != '\0'; ++k) h = h * 31u + (unsigned)(int)(signed char)p[k];
  return h & 0x7FFFFFFFu;
}

/* Prefix equality: 0 as soon as either side reaches NUL, else the first byte
 * difference.  com/cpp/cpu-runtime.hpp:37-45 */
inline int StrCmp(const char* a, const char* b, int n) {
  for (int k = 0; k < n; ++k) {
    if (a[k] == '\0' || b[k] == '\0') return 0;
    if (a[k] != b[k]) return (int)(signed char)a[k] - (int)(signed char)b[k];
  }
  return 0;
}

/* com/cpp/cpu-runtime.hpp:76-78 */
inline void StrCopy(char* dst, const char* src, int n) {
  for (int k = 0; k < n; ++k) dst[k] = src[k];
}

/* Linear probing from Hash(key bytes) % ht_size, at most 10 slots, silent drop.
 * com/cpp/cpu-runtime.hpp:80-91 */
template <typename T>
inline bool InsertHashTable(const T* tab, int pos, unsigned key_size, int* ht, unsigned ht_size) {
  unsigned slot = Hash((const char*)(tab + pos), (int)key_size) % ht_size;
  for (int t = 0; t < 10; ++t) {
    if (ht[slot] < 0) {
      ht[slot] = pos;
      return true;
    }
    slot = (slot + 1) % ht_size;
  }
  return false;
}

/* String keys hash with StrHash over the fixed-width value.
 * com/cpp/cpu-runtime.hpp:92-103 */
template <>
inline bool InsertHashTable<char>(const char* tab, int pos, unsigned key_size, int* ht,
                                  unsigned ht_size) {
  unsigned slot = StrHash(tab + (uint64)pos * key_size, (int)key_size) % ht_size;
  for (int t = 0; t < 10; ++t) {
    if (ht[slot] < 0) {
      ht[slot] = pos;
      return true;
    }
    slot = (slot + 1) % ht_size;
  }
  return false;
}

/* First match in probe order, -1 on an empty slot or after 10 slots.
 * com/cpp/cpu-runtime.hpp:125-138 */
template <typename T>
inline int Probe(const T* key, unsigned key_size, const T* tab, const int* ht, unsigned ht_size) {
  unsigned slot = Hash((const char*)key, (int)key_size) % ht_size;
  for (int t = 0; t < 10; ++t) {
    int pos = ht[slot];
    if (pos < 0) return -1;
    if (tab[pos] == *key) return pos;
    slot = (slot + 1) % ht_size;
  }
  return -1;
}

/* String variant: StrHash home slot, StrCmp (prefix) equality, and 0 (a valid row
 * id) rather than -1 after 10 slots.  com/cpp/cpu-runtime.hpp:140-152 */
template <>
inline int Probe<char>(const char* key, unsigned key_size, const char* tab, const int* ht,
                       unsigned ht_size) {
  unsigned slot = StrHash(key, (int)key_size) % ht_size;
  for (int t = 0; t < 10; ++t) {
    int pos = ht[slot];
    if (pos < 0) return -1;
    if (StrCmp(key, tab + (uint64)pos * key_size, (int)key_size) == 0) return pos;
    slot = (slot + 1) % ht_size;
  }
  return 0;
}

/* Fill with -1, insert rows in row order; returns the number of dropped inserts
 * (the reference ignores InsertHashTable's result).  com/cpp/com.hpp:168-174 */
template <typename T>
inline uint64 BuildHashTable(const T* tab, unsigned tab_sz, unsigned key_sz, int* ht,
                             unsigned ht_sz) {
  for (unsigned s = 0; s < ht_sz; ++s) ht[s] = -1;
  uint64 dropped = 0;
  for (unsigned r = 0; r < tab_sz; ++r)
    if (!InsertHashTable<T>(tab, (int)r, key_sz, ht, ht_sz)) ++dropped;
  return dropped;
}

/* ------------------------------------------------------------------------- */
/* Driver structures (plain C layout, filled from Python through ctypes).     */

#define ORT_MAX_OBJ 8
#define ORT_MAX_COL 24
#define ORT_MAX_KEY 3
#define ORT_MAX_MET 8

enum { ORT_Byte = 0, ORT_Int16 = 1, ORT_Int32 = 2, ORT_Int64 = 3, ORT_String = 7 };

struct ORT_Obj {
  const char* col[ORT_MAX_COL];
  uint64 rows;          /* tuple count (dimensions: rows of bulk 0) */
  uint64 max_tuple_ct;  /* capacity of one bulk; fixes ht_size */
  int32_t ncols;
  int32_t type[ORT_MAX_COL];
  int64_t size[ORT_MAX_COL];
};

struct ORT_Ctx {
  int32_t nobj;
  int32_t threads;
  uint64 tuple_per_bulk;
  ORT_Obj obj[ORT_MAX_OBJ];
  /* filled by the build (Pipeline::Open) */
  int* ht[ORT_MAX_OBJ];
  uint64 ht_size[ORT_MAX_OBJ];
  uint64 dropped[ORT_MAX_OBJ];
  uint64 inserted[ORT_MAX_OBJ];
  double open_ms;
  double run_ms;
  double merge_ms;
};

enum { ORT_KIND_SELECT = 0, ORT_KIND_GROUP = 1, ORT_KIND_SCALAR = 2 };

struct ORT_Result {
  int32_t kind;
  int32_t ncols;
  uint64 rows;
  char* col[ORT_MAX_COL];
  int64_t width[ORT_MAX_COL];
  uint64 bulk0_rows;  /* select: rows of sink bulk 0 = what source bulk 0 produced */
};

/* Per-bulk output written by the generated task. */
struct ORT_BulkOut {
  char* col[ORT_MAX_COL];     /* select: projected columns, compacted */
  int32_t* key[ORT_MAX_KEY];  /* group-by: key columns, compacted */
  int64_t* met[ORT_MAX_MET];  /* group-by: metric columns, compacted; scalar: met[m][0] */
  uint64 rows;
};

typedef uint64 (*ORT_TaskFn)(const ORT_Ctx* ctx, const char* const* fact, uint64 tuple_ct,
                             ORT_BulkOut* out);

static inline double ort_ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

/* Hash-table build for one join's rkey over bulk 0 of ref object `o`.
 * core/pipeline.cu:88-133 (only bulk 0; ht_size = HashRate * max_tuple_count). */
template <typename T>
inline void ort_build_join(ORT_Ctx* c, int o, int col) {
  ORT_Obj& ob = c->obj[o];
  /* BuildHashTable takes unsigned sizes (com/cpp/com.hpp:169); ob.rows of a ref
   * object is already its bulk 0 tuple count */
  uint64 ht_size = (unsigned)(HashRate * ob.max_tuple_ct);
  uint64 rows = ob.rows;
  if (c->ht[o] == nullptr) c->ht[o] = (int*)malloc(sizeof(int) * ht_size);
  c->ht_size[o] = ht_size;
  c->dropped[o] = BuildHashTable<T>((const T*)ob.col[col], (unsigned)rows,
                                    (unsigned)ob.size[col], c->ht[o], (unsigned)ht_size);
  c->inserted[o] = rows - c->dropped[o];
}

struct ORT_GroupRow {
  int32_t k[ORT_MAX_KEY];
  int64_t m[ORT_MAX_MET];
};

/* sort_by_key on the key tuple (lexicographic, signed) then reduce_by_key with
 * tuple sums, as ReduceTask does per bulk (core/pipeline.cu:380-471), in int64. */
static inline void ort_sort_reduce(std::vector<ORT_GroupRow>& rows, int nk, int nm) {
  std::sort(rows.begin(), rows.end(), [nk](const ORT_GroupRow& a, const ORT_GroupRow& b) {
    for (int q = 0; q < nk; ++q)
      if (a.k[q] != b.k[q]) return a.k[q] < b.k[q];
    return false;
  });
  size_t w = 0;
  for (size_t r = 0; r < rows.size(); ++r) {
    bool same = w > 0;
    for (int q = 0; same && q < nk; ++q) same = rows[w - 1].k[q] == rows[r].k[q];
    if (same) {
      for (int q = 0; q < nm; ++q) rows[w - 1].m[q] += rows[r].m[q];
    } else {
      rows[w++] = rows[r];
    }
  }
  rows.resize(w);
}

/*
 * Run all bulks of obj 0 with `threads` workers (the reference spawns one thread
 * per bulk behind a 16-slot CPU semaphore, core/resource.cu:21,89-105), then merge.
 * kind/nk/nm/out_width describe the output; out_width[c] is the byte width of
 * projected column c for select pipelines.
 */
static inline int ort_drive(ORT_Ctx* c, ORT_Result* res, ORT_TaskFn task, int kind, int nk,
                            int nm, int nout, const int64_t* out_width) {
  const ORT_Obj& fact = c->obj[0];
  const uint64 tpb = c->tuple_per_bulk;
  const uint64 nbulk = fact.rows == 0 ? 0 : (fact.rows + tpb - 1) / tpb;
  auto t0 = std::chrono::steady_clock::now();

  std::vector<std::vector<ORT_GroupRow>> parts(nbulk);
  std::vector<std::vector<std::vector<char>>> sel(nbulk);
  std::vector<std::vector<int64_t>> scal(nbulk, std::vector<int64_t>(ORT_MAX_MET, 0));
  std::vector<uint64> sel_rows(nbulk, 0);
  std::atomic<uint64> next(0);

  auto worker = [&]() {
    std::vector<int32_t> kbuf[ORT_MAX_KEY];
    std::vector<int64_t> mbuf[ORT_MAX_MET];
    for (;;) {
      uint64 b = next.fetch_add(1);
      if (b >= nbulk) break;
      uint64 r0 = b * tpb;
      uint64 n = std::min<uint64>(tpb, fact.rows - r0);
      const char* fcols[ORT_MAX_COL];
      for (int q = 0; q < fact.ncols; ++q)
        fcols[q] = fact.col[q] ? fact.col[q] + r0 * (uint64)fact.size[q] : nullptr;
      ORT_BulkOut out;
      memset(&out, 0, sizeof(out));
      if (kind == ORT_KIND_SELECT) {
        sel[b].resize(nout);
        for (int q = 0; q < nout; ++q) {
          sel[b][q].assign(n * (uint64)out_width[q] + 1, 0);
          out.col[q] = sel[b][q].data();
        }
      } else if (kind == ORT_KIND_GROUP) {
        for (int q = 0; q < nk; ++q) { kbuf[q].resize(n + 1); out.key[q] = kbuf[q].data(); }
        for (int q = 0; q < nm; ++q) { mbuf[q].resize(n + 1); out.met[q] = mbuf[q].data(); }
      } else {
        for (int q = 0; q < nm; ++q) out.met[q] = &scal[b][q];
      }
      uint64 j = task(c, fcols, n, &out);
      if (kind == ORT_KIND_SELECT) {
        sel_rows[b] = j;
      } else if (kind == ORT_KIND_GROUP) {
        std::vector<ORT_GroupRow>& p = parts[b];
        p.resize(j);
        for (uint64 r = 0; r < j; ++r) {
          for (int q = 0; q < nk; ++q) p[r].k[q] = out.key[q][r];
          for (int q = 0; q < nm; ++q) p[r].m[q] = out.met[q][r];
        }
        ort_sort_reduce(p, nk, nm);
      }
    }
  };
  int nt = c->threads > 0 ? c->threads : 1;
  if ((uint64)nt > nbulk) nt = (int)(nbulk > 0 ? nbulk : 1);
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  c->run_ms = ort_ms_since(t0);

  auto t1 = std::chrono::steady_clock::now();
  memset(res, 0, sizeof(*res));
  res->kind = kind;
  if (kind == ORT_KIND_SELECT) {
    uint64 total = 0;
    for (uint64 b = 0; b < nbulk; ++b) total += sel_rows[b];
    res->ncols = nout;
    res->rows = total;
    /* sink bulk b holds what source bulk b produced (core/pipeline.cu:161-162,294-296) */
    res->bulk0_rows = nbulk ? sel_rows[0] : 0;
    for (int q = 0; q < nout; ++q) {
      res->width[q] = out_width[q];
      res->col[q] = (char*)malloc(total * (uint64)out_width[q] + 1);
      uint64 off = 0;
      for (uint64 b = 0; b < nbulk; ++b) {
        uint64 bytes = sel_rows[b] * (uint64)out_width[q];
        if (bytes) memcpy(res->col[q] + off, sel[b][q].data(), bytes);
        off += bytes;
      }
    }
  } else if (kind == ORT_KIND_GROUP) {
    std::vector<ORT_GroupRow> all;
    for (uint64 b = 0; b < nbulk; ++b) all.insert(all.end(), parts[b].begin(), parts[b].end());
    ort_sort_reduce(all, nk, nm);
    res->ncols = nk + nm;
    res->rows = all.size();
    for (int q = 0; q < nk; ++q) {
      res->width[q] = 4;
      int32_t* d = (int32_t*)malloc(4 * all.size() + 4);
      for (size_t r = 0; r < all.size(); ++r) d[r] = all[r].k[q];
      res->col[q] = (char*)d;
    }
    for (int q = 0; q < nm; ++q) {
      res->width[nk + q] = 8;
      int64_t* d = (int64_t*)malloc(8 * all.size() + 8);
      for (size_t r = 0; r < all.size(); ++r) d[r] = all[r].m[q];
      res->col[nk + q] = (char*)d;
    }
  } else {
    /* The scalar query yields exactly one row (core/pipeline.cu:373-379). */
    res->ncols = nm;
    res->rows = 1;
    for (int q = 0; q < nm; ++q) {
      int64_t s = 0;
      for (uint64 b = 0; b < nbulk; ++b) s += scal[b][q];
      res->width[q] = 8;
      int64_t* d = (int64_t*)malloc(8);
      *d = s;
      res->col[q] = (char*)d;
    }
  }
  c->merge_ms = ort_ms_since(t1);
  return 0;
}

static inline void ort_free_result(ORT_Result* r) {
  for (int q = 0; q < ORT_MAX_COL; ++q) {
    free(r->col[q]);
    r->col[q] = nullptr;
  }
}

static inline void ort_free_ctx(ORT_Ctx* c) {
  for (int o = 0; o < ORT_MAX_OBJ; ++o) {
    free(c->ht[o]);
    c->ht[o] = nullptr;
  }
}

#endif /* SVM_ORACLE_RT_H_ */
