// datagen_host.cpp — host side of the synthetic star-schema generator: the date
// dimension, Zipf CDF tables, multi-threaded host generation and device tables.
#include <cmath>
#include <cstring>
#include <thread>

#include "datagen.h"
#include "kernels.h"
#include "svm_internal.h"

namespace svm {

namespace {

struct ZipfTable {
  std::vector<uint64_t> cdf;
  std::vector<uint32_t> guide;
  uint32_t bits = 0;
  DevBuf dcdf, dguide;
};

std::mutex g_zipf_mu;
std::map<std::pair<uint64_t, double>, std::shared_ptr<ZipfTable>> g_zipf;

// cdf[i] = floor(2^64 * P(rank <= i+1)) for P(rank = k) proportional to k^-s;
// guide[g] = first i with cdf[i] >= g * 2^(64-bits).
std::shared_ptr<ZipfTable> zipf_table(uint64_t n, double s) {
  std::lock_guard<std::mutex> g(g_zipf_mu);
  auto key = std::make_pair(n, s);
  auto it = g_zipf.find(key);
  if (it != g_zipf.end()) return it->second;
  auto z = std::make_shared<ZipfTable>();
  z->cdf.resize(n);
  long double total = 0;
  for (uint64_t k = 1; k <= n; ++k) total += powl((long double)k, -(long double)s);
  long double acc = 0;
  const long double two64 = 18446744073709551616.0L;
  for (uint64_t k = 1; k <= n; ++k) {
    acc += powl((long double)k, -(long double)s);
    long double v = acc / total * two64;
    z->cdf[k - 1] = v >= two64 ? UINT64_MAX : (uint64_t)v;
  }
  z->cdf[n - 1] = UINT64_MAX;
  uint32_t bits = 1;
  while (bits < 16 && (1ull << bits) < n) bits++;
  z->bits = bits;
  z->guide.resize((1ull << bits) + 1);
  uint64_t i = 0;
  for (uint64_t gi = 0; gi < (1ull << bits); ++gi) {
    uint64_t th = gi << (64 - bits);
    while (i < n - 1 && z->cdf[i] < th) i++;
    z->guide[gi] = (uint32_t)i;
  }
  z->guide[1ull << bits] = (uint32_t)(n - 1);
  g_zipf[key] = z;
  return z;
}

const std::vector<int32_t>& date_rows() {
  // SVM_DATE_ROWS consecutive days from 1992-01-01: datekey yyyymmdd, year, month.
  static std::vector<int32_t> rows = [] {
    std::vector<int32_t> v;
    int y = 1992, m = 1, d = 1;
    auto dim = [](int yy, int mm) {
      static const int days[] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
      bool leap = (yy % 4 == 0 && yy % 100 != 0) || yy % 400 == 0;
      return mm == 2 && leap ? 29 : days[mm - 1];
    };
    for (int r = 0; r < SVM_DATE_ROWS; ++r) {
      v.push_back(y * 10000 + m * 100 + d);
      v.push_back(y);
      v.push_back(m);
      if (++d > dim(y, m)) {
        d = 1;
        if (++m > 12) { m = 1; y++; }
      }
    }
    return v;
  }();
  return rows;
}

std::vector<int32_t> datekeys() {
  std::vector<int32_t> k;
  const auto& r = date_rows();
  for (size_t i = 0; i < r.size(); i += 3) k.push_back(r[i]);
  return k;
}

void fill_zipf(svm_zipf& z, const ZipfTable& t, uint64_t n, bool device) {
  z.cdf = device ? t.dcdf.as<uint64_t>() : t.cdf.data();
  z.guide = device ? t.dguide.as<uint32_t>() : t.guide.data();
  z.n = n;
  z.guide_bits = t.bits;
  z.perm_mul = 1000003ull;
}

void ensure_device(ZipfTable& t) {
  std::lock_guard<std::mutex> g(g_zipf_mu);
  if (t.dcdf.p) return;
  DevBuf c(t.cdf.size() * 8), gd(t.guide.size() * 4);
  SVM_HIP(hipMemcpy(c.p, t.cdf.data(), t.cdf.size() * 8, hipMemcpyHostToDevice));
  SVM_HIP(hipMemcpy(gd.p, t.guide.data(), t.guide.size() * 4, hipMemcpyHostToDevice));
  t.dcdf = std::move(c);
  t.dguide = std::move(gd);
}

}  // namespace

std::vector<std::pair<int, int64_t>> gen_schema(int kind) {
  std::vector<std::pair<int, int64_t>> s;
  auto ints = [&](int n) { for (int q = 0; q < n; ++q) s.push_back({Int32, 4}); };
  switch (kind) {
    case SVM_GEN_DATE: ints(3); break;
    case SVM_GEN_FACT1: ints(4); break;
    case SVM_GEN_PART2: case SVM_GEN_PART3: ints(3); break;
    case SVM_GEN_SUPP2: ints(4); break;
    case SVM_GEN_FACT2: ints(4); break;
    case SVM_GEN_CUST3: case SVM_GEN_SUPP3:
      ints(1);
      for (int q = 0; q < 3; ++q) s.push_back({String, SVM_GEN_STR});
      break;
    case SVM_GEN_FACT3: ints(6); break;
    default: fail(SVM_ERR_INVALID, "unknown generator kind " + std::to_string(kind));
  }
  return s;
}

struct GenState {
  svm_gen_ctx ctx;
  std::vector<int32_t> keys;
  std::shared_ptr<ZipfTable> z[4];
  DevBuf dkeys;
};

static void gen_prepare(GenState& st, int kind, uint64_t seed, const svm_gen_params* p, bool device) {
  memset(&st.ctx, 0, sizeof(st.ctx));
  st.ctx.seed = seed;
  st.keys = datekeys();
  st.ctx.n_dates = st.keys.size();
  st.ctx.n_cust = p ? p->n_cust : 0;
  st.ctx.n_supp = p ? p->n_supp : 0;
  st.ctx.n_part = p ? p->n_part : 0;
  if (device) {
    st.dkeys = DevBuf(st.keys.size() * 4);
    SVM_HIP(hipMemcpy(st.dkeys.p, st.keys.data(), st.keys.size() * 4, hipMemcpyHostToDevice));
    st.ctx.datekeys = st.dkeys.as<int32_t>();
  } else {
    st.ctx.datekeys = st.keys.data();
  }
  if (kind == SVM_GEN_FACT3) {
    double s = p && p->zipf_s > 0 ? p->zipf_s : 1.1;
    uint64_t ns[4] = {st.ctx.n_dates, st.ctx.n_cust, st.ctx.n_supp, st.ctx.n_part};
    for (int q = 0; q < 4; ++q) {
      if (ns[q] == 0) fail(SVM_ERR_INVALID, "FACT3 needs n_cust, n_supp and n_part");
      st.z[q] = zipf_table(ns[q], s);
      if (device) ensure_device(*st.z[q]);
      fill_zipf(st.ctx.zipf[q], *st.z[q], ns[q], device);
    }
  }
  if ((kind == SVM_GEN_FACT2) && (st.ctx.n_part == 0 || st.ctx.n_supp == 0))
    fail(SVM_ERR_INVALID, "FACT2 needs n_supp and n_part");
}

void gen_host(int kind, uint64_t seed, const svm_gen_params* p, uint64_t row0, uint64_t n,
              void* const* cols, int threads) {
  auto schema = gen_schema(kind);
  if (kind == SVM_GEN_DATE) {
    const auto& r = date_rows();
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t row = row0 + i;
      if (row >= SVM_DATE_ROWS) fail(SVM_ERR_INVALID, "the date dimension has 2556 rows");
      for (int c = 0; c < 3; ++c) ((int32_t*)cols[c])[i] = r[row * 3 + c];
    }
    return;
  }
  GenState st;
  gen_prepare(st, kind, seed, p, false);
  int T = threads > 0 ? threads : 1;
  std::vector<std::thread> pool;
  auto work = [&](int t) {
    uint64_t b = n * (uint64_t)t / T, e = n * (uint64_t)(t + 1) / T;
    void* shifted[SVM_GEN_MAX_COLS] = {nullptr};
    for (size_t c = 0; c < schema.size(); ++c)
      shifted[c] = (char*)cols[c] + b * (uint64_t)schema[c].second;
    for (uint64_t i = b; i < e; ++i) svm_gen_row(&st.ctx, kind, row0 + i, i - b, shifted);
  };
  for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
}

TablePtr gen_device_table(const std::string& id, int kind, uint64_t rows, uint64_t seed,
                          const svm_gen_params* p) {
  Engine& E = engine();
  auto schema = gen_schema(kind);
  if (kind == SVM_GEN_DATE && rows > SVM_DATE_ROWS) fail(SVM_ERR_INVALID, "the date dimension has 2556 rows");
  TablePtr t = E.store.create(id, schema, rows);
  t->max_tuple_count = E.cfg.tuple_per_bulk;
  if (kind == SVM_GEN_DATE) {
    const auto& r = date_rows();
    for (int c = 0; c < 3; ++c) {
      std::vector<int32_t> col(rows);
      for (uint64_t i = 0; i < rows; ++i) col[i] = r[i * 3 + c];
      SVM_HIP(hipMemcpy(t->cols[c].dev, col.data(), rows * 4, hipMemcpyHostToDevice));
    }
  } else {
    GenState st;
    gen_prepare(st, kind, seed, p, true);
    void* cols[SVM_GEN_MAX_COLS] = {nullptr};
    for (size_t c = 0; c < schema.size(); ++c) cols[c] = t->cols[c].dev;
    k::gen_table(st.ctx, kind, 0, rows, cols, E.stream);
    SVM_HIP(hipStreamSynchronize(E.stream));
  }
  E.store.put(id, t);
  return t;
}

}  // namespace svm
