/*
 * datagen.h — seeded synthetic star-schema generator, identical on host and device.
 *
 * Every value is a pure function of (seed, table kind, column stream, row), so the
 * host (for the CPU oracle and the CPU baseline) and the GPU (for full-size device
 * resident tables) produce byte-identical columns without copying data around.
 *
 * Random bits: the splitmix64 output function (Steele, Lea, Flood, "Fast splittable
 * pseudorandom number generators", OOPSLA 2014) applied to a per-(seed, stream)
 * key plus the row index.  Bounded integers use the multiply-shift reduction
 * (Lemire 2019) on the high 32 bits, which is deterministic and branch-free.
 *
 * Zipf(s) foreign keys use inverse-CDF sampling over an integer CDF table
 * (uint64 thresholds, computed once on the host and shared by both sides), with a
 * guide table (Chen & Asau 1974) to bound the binary search.  Rank r is scattered
 * to key ((r-1)*perm_mul mod n)+1 so popular keys are not adjacent.
 *
 * Table shapes (column order is the table's schema order):
 *   DATE   : datekey(yyyymmdd, consecutive days from 1992-01-01), year, month   [host only]
 *   FACT1  : orderdate, quantity[1,50], discount[0,10], extendedprice[1000,1e7]
 *   PART2  : partkey(1..N), category[1,25], brand[(cat-1)*40+1 .. cat*40]
 *   SUPP2  : suppkey(1..N), region, nation, city           (int codes)
 *   FACT2  : orderdate, partkey, suppkey, revenue[1000,1e7] (uniform FKs)
 *   CUST3  : custkey(1..N), city, nation, region           (10-byte NUL padded strings)
 *   SUPP3  : suppkey(1..N), city, nation, region           (10-byte strings)
 *   PART3  : partkey(1..N), category, brand
 *   FACT3  : orderdate, custkey, suppkey, partkey, revenue[1000,1e7], supplycost[1000,120000]
 *            (all four foreign keys Zipf-skewed)
 * Geography: nation uniform over 25, region = nation/5, city = region*50 + U[0,49]
 * (250 cities, 50 per region, independent of the nation inside its region).
 */
#ifndef SVM_DATAGEN_H_
#define SVM_DATAGEN_H_

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define SVM_HD __host__ __device__ inline
#else
#define SVM_HD static inline
#endif

enum svm_gen_kind {
  SVM_GEN_DATE = 0,
  SVM_GEN_FACT1 = 1,
  SVM_GEN_PART2 = 2,
  SVM_GEN_SUPP2 = 3,
  SVM_GEN_FACT2 = 4,
  SVM_GEN_CUST3 = 5,
  SVM_GEN_SUPP3 = 6,
  SVM_GEN_PART3 = 7,
  SVM_GEN_FACT3 = 8,
  SVM_GEN_KIND_COUNT = 9
};

#define SVM_GEN_STR 10          /* fixed width of generated string columns */
#define SVM_GEN_MAX_COLS 6
#define SVM_DATE_ROWS 2556

typedef struct svm_zipf {
  const uint64_t* cdf;    /* cdf[i] = floor(P(rank <= i+1) * 2^64), cdf[n-1] = 2^64-1 */
  const uint32_t* guide;  /* (1<<guide_bits)+1 entries */
  uint64_t n;
  uint32_t guide_bits;
  uint32_t pad_;
  uint64_t perm_mul;      /* prime not dividing n */
} svm_zipf;

typedef struct svm_gen_ctx {
  uint64_t seed;
  const int32_t* datekeys; /* SVM_DATE_ROWS consecutive yyyymmdd keys */
  uint64_t n_dates;
  uint64_t n_cust, n_supp, n_part; /* foreign-key ranges of the fact tables */
  svm_zipf zipf[4];                /* FACT3: orderdate index, custkey, suppkey, partkey */
} svm_gen_ctx;

SVM_HD uint64_t svm_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* 64 random bits for (seed, stream, row). */
SVM_HD uint64_t svm_rand(uint64_t seed, uint32_t stream, uint64_t row) {
  uint64_t key = svm_mix64(seed ^ ((uint64_t)stream * 0xD1342543DE82EF95ull));
  return svm_mix64(key + row);
}

/* Uniform integer in [lo, hi] (hi - lo < 2^32). */
SVM_HD int32_t svm_uniform(uint64_t r, int64_t lo, int64_t hi) {
  uint64_t span = (uint64_t)(hi - lo + 1);
  return (int32_t)(lo + (int64_t)(((r >> 32) * span) >> 32));
}

/* Zipf rank in [1, n] for 64 random bits u. */
SVM_HD uint64_t svm_zipf_rank(const svm_zipf* z, uint64_t u) {
  uint64_t g = u >> (64 - z->guide_bits);
  uint64_t lo = z->guide[g];
  uint64_t hi = z->guide[g + 1];
  while (lo < hi) { /* first i in [lo, hi] with cdf[i] >= u */
    uint64_t mid = lo + ((hi - lo) >> 1);
    if (z->cdf[mid] >= u) hi = mid; else lo = mid + 1;
  }
  return lo + 1;
}

SVM_HD uint64_t svm_zipf_key(const svm_zipf* z, uint64_t u) {
  uint64_t r = svm_zipf_rank(z, u);
  return ((r - 1) * z->perm_mul) % z->n + 1;
}

SVM_HD void svm_put_name(char* dst, const char* prefix, int plen, int value, int digits) {
  int k = 0;
  for (; k < plen; k++) dst[k] = prefix[k];
  int p10 = 1;
  for (int d = 1; d < digits; d++) p10 *= 10;
  for (int d = 0; d < digits; d++) {
    dst[k++] = (char)('0' + (value / p10) % 10);
    p10 /= 10;
  }
  for (; k < SVM_GEN_STR; k++) dst[k] = '\0';
}

/* Geography codes shared by SUPP2/CUST3/SUPP3: nation U[0,24], region nation/5,
 * city region*50 + U[0,49]. */
SVM_HD void svm_geo(uint64_t seed, uint32_t kind, uint64_t row, int32_t* region,
                    int32_t* nation, int32_t* city) {
  int32_t n = svm_uniform(svm_rand(seed, kind * 16u + 1u, row), 0, 24);
  int32_t r = n / 5;
  int32_t c = r * 50 + svm_uniform(svm_rand(seed, kind * 16u + 2u, row), 0, 49);
  *region = r;
  *nation = n;
  *city = c;
}

/*
 * Generate one row of table `kind` into the column arrays `cols` (each column
 * addressed as cols[c] + row_local * width(c)).  Int columns are int32,
 * string columns SVM_GEN_STR bytes.
 */
SVM_HD void svm_gen_row(const svm_gen_ctx* ctx, int kind, uint64_t row, uint64_t out_row,
                        void* const* cols) {
  const uint64_t s = ctx->seed;
  const uint32_t k = (uint32_t)kind * 16u;
  switch (kind) {
    case SVM_GEN_FACT1: {
      int32_t* c0 = (int32_t*)cols[0];
      int32_t* c1 = (int32_t*)cols[1];
      int32_t* c2 = (int32_t*)cols[2];
      int32_t* c3 = (int32_t*)cols[3];
      uint32_t di = (uint32_t)svm_uniform(svm_rand(s, k + 1u, row), 0, (int64_t)ctx->n_dates - 1);
      c0[out_row] = ctx->datekeys[di];
      c1[out_row] = svm_uniform(svm_rand(s, k + 2u, row), 1, 50);
      c2[out_row] = svm_uniform(svm_rand(s, k + 3u, row), 0, 10);
      c3[out_row] = svm_uniform(svm_rand(s, k + 4u, row), 1000, 10000000);
      break;
    }
    case SVM_GEN_PART2:
    case SVM_GEN_PART3: {
      int32_t* c0 = (int32_t*)cols[0];
      int32_t* c1 = (int32_t*)cols[1];
      int32_t* c2 = (int32_t*)cols[2];
      int32_t cat = svm_uniform(svm_rand(s, k + 1u, row), 1, 25);
      c0[out_row] = (int32_t)(row + 1);
      c1[out_row] = cat;
      c2[out_row] = (cat - 1) * 40 + svm_uniform(svm_rand(s, k + 2u, row), 1, 40);
      break;
    }
    case SVM_GEN_SUPP2: {
      int32_t r, n, c;
      svm_geo(s, (uint32_t)kind, row, &r, &n, &c);
      ((int32_t*)cols[0])[out_row] = (int32_t)(row + 1);
      ((int32_t*)cols[1])[out_row] = r;
      ((int32_t*)cols[2])[out_row] = n;
      ((int32_t*)cols[3])[out_row] = c;
      break;
    }
    case SVM_GEN_FACT2: {
      uint32_t di = (uint32_t)svm_uniform(svm_rand(s, k + 1u, row), 0, (int64_t)ctx->n_dates - 1);
      ((int32_t*)cols[0])[out_row] = ctx->datekeys[di];
      ((int32_t*)cols[1])[out_row] = svm_uniform(svm_rand(s, k + 2u, row), 1, (int64_t)ctx->n_part);
      ((int32_t*)cols[2])[out_row] = svm_uniform(svm_rand(s, k + 3u, row), 1, (int64_t)ctx->n_supp);
      ((int32_t*)cols[3])[out_row] = svm_uniform(svm_rand(s, k + 4u, row), 1000, 10000000);
      break;
    }
    case SVM_GEN_CUST3:
    case SVM_GEN_SUPP3: {
      int32_t r, n, c;
      svm_geo(s, (uint32_t)kind, row, &r, &n, &c);
      ((int32_t*)cols[0])[out_row] = (int32_t)(row + 1);
      svm_put_name((char*)cols[1] + out_row * SVM_GEN_STR, "CITY_", 5, c, 3);
      svm_put_name((char*)cols[2] + out_row * SVM_GEN_STR, "NATION_", 7, n, 2);
      svm_put_name((char*)cols[3] + out_row * SVM_GEN_STR, "REGION_", 7, r, 1);
      break;
    }
    case SVM_GEN_FACT3: {
      uint64_t di = svm_zipf_key(&ctx->zipf[0], svm_rand(s, k + 1u, row)) - 1;
      ((int32_t*)cols[0])[out_row] = ctx->datekeys[di];
      ((int32_t*)cols[1])[out_row] = (int32_t)svm_zipf_key(&ctx->zipf[1], svm_rand(s, k + 2u, row));
      ((int32_t*)cols[2])[out_row] = (int32_t)svm_zipf_key(&ctx->zipf[2], svm_rand(s, k + 3u, row));
      ((int32_t*)cols[3])[out_row] = (int32_t)svm_zipf_key(&ctx->zipf[3], svm_rand(s, k + 4u, row));
      ((int32_t*)cols[4])[out_row] = svm_uniform(svm_rand(s, k + 5u, row), 1000, 10000000);
      ((int32_t*)cols[5])[out_row] = svm_uniform(svm_rand(s, k + 6u, row), 1000, 120000);
      break;
    }
    default:
      break;
  }
}

#endif /* SVM_DATAGEN_H_ */
