// batch_kernels.hip — parameter-driven kernels of the batched path: dimension
// profiles and their per-batch query masks, the shared scan in which every fact
// row is read once for all queries of the batch, and batched ordered group emission.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "batch.h"
#include "svm_internal.h"

namespace svm {
namespace bk {

namespace {

constexpr int kBlock = 256;
typedef int i32x4 __attribute__((ext_vector_type(4)));

// The reference's StrCmp: compare up to n bytes, stopping at the first NUL of
// either side (define.hpp StrCmp).
__device__ __forceinline__ int ref_strcmp(const char* a, const char* b, int n) {
  for (int k = 0; k < n; ++k) {
    char x = a[k], y = b[k];
    if (x == '\0' || y == '\0') return 0;
    if (x != y) return (int)(signed char)x - (int)(signed char)y;
  }
  return 0;
}

// ------------------------------------------------------------------ profiles
__global__ void k_profile_raw(RawArgs a, uint64_t n, int32_t* __restrict__ raw) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    unsigned v = 0;
    for (int k = 0; k < a.nattr; ++k) v |= (unsigned)a.codes[k][r] << a.shift[k];
    raw[r] = (int32_t)v;
  }
}

__global__ void k_profile_scatter(const int32_t* __restrict__ keys, uint64_t n, int32_t kmin,
                                  const int32_t* __restrict__ pcode, void* __restrict__ pkey, int narrow) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned off = (unsigned)(keys[r] - kmin);
    if (narrow) ((uint8_t*)pkey)[off] = (uint8_t)pcode[r];
    else ((uint16_t*)pkey)[off] = (uint16_t)pcode[r];
  }
}

// Profile of key offset k of a slot; P when no dimension row has the key.
__device__ __forceinline__ unsigned slot_profile(const Slot& S, unsigned k) {
  return S.narrow ? (unsigned)((const uint8_t*)S.pkey)[k] : (unsigned)((const uint16_t*)S.pkey)[k];
}

// Per batch and slot, one bit per key offset: the key has a dimension row whose
// profile passes some query that joins the slot.  One wave per 64 keys.
__global__ void k_slot_bits(const ScanArgs* __restrict__ A) {
  const unsigned lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (int s = 0; s < A->nslot; ++s) {
    const Slot& S = A->slot[s];
    const uint64_t chunks = ((uint64_t)S.range + 63) / 64;
    for (uint64_t c = wave; c < chunks; c += nwaves) {
      const uint64_t k = c * 64 + lane;
      bool pass = false;
      if (k < S.range) {
        const unsigned p = slot_profile(S, (unsigned)k);
        pass = p < S.P && (S.mask[p] & S.joined) != 0ull;
      }
      const unsigned long long b = __ballot(pass);
      if (lane == 0) {
        S.bits[2 * c] = (unsigned)b;
        S.bits[2 * c + 1] = (unsigned)(b >> 32);
      }
    }
  }
}

__global__ void k_string_exact(const char* __restrict__ col, int size, uint64_t n, const int32_t* __restrict__ codes,
                               const int32_t* __restrict__ rep, unsigned* flag) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    const char* a = col + r * (uint64_t)size;
    const char* b = col + (uint64_t)rep[codes[r]] * (uint64_t)size;
    bool same = true;
    for (int k = 0; k < size; ++k) same = same && a[k] == b[k];
    if (!same) atomicOr(flag, 1u);
  }
}

// One thread per (slot, profile, run of kTermsPerThread terms), ORing its queries
// into the profile's mask (cleared by k_pass_clear); the grid covers kMaxSlot x
// kMaxProfiles x kMaxTerms so the launch does not depend on the batch.
constexpr int kTermsPerThread = 8;
__global__ void k_profile_masks(const MaskArgs* __restrict__ M) {
  if ((int)blockIdx.y >= M->nslot) return;
  const MaskSlot& S = M->slots[blockIdx.y];
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  const int t0 = S.t0 + (int)blockIdx.z * kTermsPerThread;
  const int t1 = min(S.t1, t0 + kTermsPerThread);
  if (p >= S.P || t0 >= t1) return;
  const unsigned raw = (unsigned)S.praw[p];
  unsigned long long m = 0ull;
  for (int t = t0; t < t1; ++t) {
    const MaskTerm& T = M->terms[t];
    bool ok = true;
    for (int k = 0; k < T.natom && ok; ++k) {
      const MaskAtom& a = T.atom[k];
      const unsigned code = a.bits ? (raw >> a.shift) & ((1u << a.bits) - 1u) : 0u;
      if (a.kind == DP_RANGE) {
        const int v = a.dict[code];
        ok = v >= a.lo && v <= a.hi;
      } else {
        const char* str = a.col + (uint64_t)a.rep[code] * (uint64_t)a.size;
        const int c = ref_strcmp(str, a.lit, a.n);
        ok = a.kind == DP_STREQ ? c == 0 : c != 0;
      }
    }
    if (ok) m |= 1ull << T.q;
  }
  if (m) atomicOr(S.mask + p, m);
}

// --------------------------------------------------------------- shared scan
// Value of fact column c (uniform across the wave) for element e of this lane: an
// indexed register read (s_set_gpr_idx), one VALU move instead of a select chain.
__device__ __forceinline__ int pick(const i32x4 (&x)[kMaxCols], int c, int e) { return x[c][e]; }

// a[i] for a lane-varying i < n, as a chain of selects kept in registers (the
// empty asm hides that the operands come from one array, which would otherwise be
// turned into an indexed load from scratch memory).
template <int N>
__device__ __forceinline__ unsigned sel(const unsigned (&a)[N], unsigned i, int n) {
  unsigned r = a[0];
#pragma unroll
  for (int k = 1; k < N; ++k) {
    if (k < n) {
      unsigned c = a[k];
      asm volatile("" : "+v"(c));
      r = i == (unsigned)k ? c : r;
    }
  }
  return r;
}

__device__ __forceinline__ int sel_i(const int (&a)[kMaxMetric], int i, int n) {
  int r = a[0];
#pragma unroll
  for (int k = 1; k < kMaxMetric; ++k) {
    if (k < n) {
      int c = a[k];
      asm volatile("" : "+v"(c));
      r = i == k ? c : r;
    }
  }
  return r;
}

// Query masks of the shared scan: 32 bits when a scan holds at most 32 queries
// (half the mask arithmetic per row), else 64.
__device__ __forceinline__ int first_bit(unsigned b) { return __ffs(b) - 1; }
__device__ __forceinline__ int first_bit(unsigned long long b) { return __ffsll((long long)b) - 1; }
template <typename M>
__device__ __forceinline__ M lds_mask(const unsigned long long* pm, unsigned i) {
  if (sizeof(M) == 4) return (M)((const unsigned*)pm)[2 * i];  // low half
  return (M)pm[i];
}

// Profiles of a slot for the four rows of a lane.  A key without a dimension row
// maps to profile P, whose mask is 0; rows that no joining query wants read entry
// `range`, which holds P as well, so the four loads issue together unconditionally.
template <typename M, bool LMAP>
__device__ __forceinline__ void slot_lookup(const Slot& S, const i32x4 (&x)[kMaxCols], const M (&gate)[4],
                                            bool use_bits, const unsigned* lmaps, unsigned (&c)[4]) {
  const unsigned range = S.range;
  const M joined = (M)S.joined;
  unsigned idx[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const unsigned off = (unsigned)(pick(x, S.fcol, e) - S.kmin);
    idx[e] = (gate[e] & joined) != (M)0 && off < range ? off : range;
  }
  if (LMAP && S.lbits >= 0) {
    const unsigned* b = lmaps + S.lbits;
#pragma unroll
    for (int e = 0; e < 4; ++e) idx[e] = (b[idx[e] >> 5] >> (idx[e] & 31u)) & 1u ? idx[e] : range;
  } else if (use_bits) {
    unsigned w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = S.bits[idx[e] >> 5];
#pragma unroll
    for (int e = 0; e < 4; ++e) idx[e] = (w[e] >> (idx[e] & 31u)) & 1u ? idx[e] : range;
  }
  if (LMAP && S.lmap >= 0) {
    const unsigned* m = lmaps + S.lmap;
    const unsigned lg = (unsigned)S.lmap_log2, per = 32u >> lg, emask = (1u << (1u << lg)) - 1u;
#pragma unroll
    for (int e = 0; e < 4; ++e) c[e] = (m[idx[e] >> (5u - lg)] >> ((idx[e] & (per - 1u)) << lg)) & emask;
  } else if (S.narrow) {
    const uint8_t* pk = (const uint8_t*)S.pkey;
#pragma unroll
    for (int e = 0; e < 4; ++e) c[e] = pk[idx[e]];
  } else {
    const uint16_t* pk = (const uint16_t*)S.pkey;
#pragma unroll
    for (int e = 0; e < 4; ++e) c[e] = pk[idx[e]];
  }
}

// Drops from each row's query mask the joining queries whose filter its profile fails.
template <typename M>
__device__ __forceinline__ void slot_apply(const Slot& S, const unsigned long long* pmask, const unsigned (&c)[4],
                                           M (&pass)[4]) {
  const M notj = ~(M)S.joined;
#pragma unroll
  for (int e = 0; e < 4; ++e) pass[e] &= lds_mask<M>(pmask, (unsigned)S.lds + c[e]) | notj;
}

// 128-bit group v of every fact column in colmask (zeros for the others).
__device__ __forceinline__ void load_groups(const ScanArgs* __restrict__ A, unsigned colmask, long long v,
                                            i32x4 (&x)[kMaxCols]) {
#pragma unroll
  for (int c = 0; c < kMaxCols; ++c) {
    if (colmask & (1u << c)) x[c] = __builtin_nontemporal_load((const i32x4*)A->cols[c] + v);
    else x[c] = i32x4{0, 0, 0, 0};
  }
}

// A wave's queue of rows to aggregate: 64 entries of query mask (2 words), profile ids
// (3 words, two slots each) and metric values (kMaxMetric words), stored field-major.
constexpr unsigned kQWords = 5 + kMaxMetric;
// Rows of a wave's group that passing rows must reach to be aggregated in place.
constexpr unsigned kQueueInline = 32;

// Group v of the columns, or zeros when v is past the tile's last group.
__device__ __forceinline__ void load_or_zero(const ScanArgs* __restrict__ A, unsigned colmask, long long v,
                                             long long vend, i32x4 (&x)[kMaxCols]) {
  if (v < vend) {
    load_groups(A, colmask, v, x);
  } else {
#pragma unroll
    for (int c = 0; c < kMaxCols; ++c) x[c] = i32x4{0, 0, 0, 0};
  }
}

// Work item: a tile of the base table.  Each 128-bit group of the columns any
// query reads is loaded once; per row, the slice, fact-predicate and slot lookups
// produce the mask of queries the row passes.  Aggregation:
//   AGG 0: each lane walks its own (row, query) pairs into one LDS hash table keyed
//          by (query, group) that overflows to the global group tables;
//   AGG 1: each lane walks the queries any of its four rows passes; the rows of one
//          group are added up in registers, then one LDS add into the query's dense
//          per-workgroup table (queries given one by the host) or one add straight
//          into its HBM table (the others): no probing, no update waits for another.
//   Q (with AGG 1): a wave whose rows pass few (query, row) tests in a group appends
//          those rows (query mask, profile ids, metric values) to its LDS queue and
//          aggregates the queue 64 rows at a time, one row per lane, so selective
//          batches do not stall the column stream on a wave's one passing row.
//   ST: staged loads.  Only the columns in A->tier0 are loaded for every group; the
//          slots 0..nslot0-1 are tested on them, and a lane loads the other columns
//          only when one of its rows still passes some query (then the fact
//          predicates and the remaining slots).  The masks are ANDed, so the order
//          of the tests does not change any result.
template <int BLOCK, int HBITS, bool PAR, bool PF, typename M, int AGG, bool LMAP, bool Q, bool ST>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) k_shared_scan(
    const ScanArgs* __restrict__ A) {
  static_assert(!(ST && (PF || PAR)), "staged loads take gated lookups without prefetch");
  constexpr unsigned kSlots = 1u << HBITS;  // LDS (query, group) accumulator
  // hash: one key word per slot; dense: one existence bit per slot
  constexpr bool DENSE = AGG != 0;
  constexpr unsigned kKeyWords = DENSE ? kSlots / 32u : kSlots;
  extern __shared__ unsigned long long lds64[];
  __shared__ unsigned long long tmask[2];
  __shared__ unsigned cur, used, tcols;
  const int nq = A->nq, nslot = A->nslot, nmetric = A->nmetric;
  const bool use_bits = A->use_bits != 0;
  const unsigned nprof = A->nprof;
  unsigned long long* hsum = lds64;
  unsigned long long* pmask = hsum + kSlots;
  unsigned* hkey = (unsigned*)(pmask + nprof);
  unsigned* praw = hkey + kKeyWords;
  QDesc* qd = (QDesc*)(praw + ((nprof + 3u) & ~3u));
  unsigned long long* lsegm = (unsigned long long*)(qd + kMaxBatch);
  int* lsegb = (int*)(lsegm + kMaxSegBounds + kMaxCols);
  unsigned* const wq = (unsigned*)(lsegb + kMaxSegBounds) + (Q ? (threadIdx.x >> 6) * (kQWords * 64u) : 0u);
  unsigned* lmaps = (unsigned*)(lsegb + kMaxSegBounds) + (Q ? (BLOCK / 64) * (kQWords * 64u) : 0u);
  unsigned qn = 0u;  // rows in the wave's queue (the same on every lane)
  unsigned long long* const gsum = A->sums;
  unsigned* const gflag = A->flags;

  // each slot's P profiles, then its no-row profile P (mask 0)
  for (int s = 0; s < nslot; ++s) {
    const unsigned P = A->slot[s].P, off = (unsigned)A->slot[s].lds;
    const unsigned long long* m = A->slot[s].mask;
    const int32_t* r = A->slot[s].praw;
    for (unsigned p = threadIdx.x; p <= P; p += BLOCK) {
      pmask[off + p] = p < P ? m[p] : 0ull;
      praw[off + p] = p < P ? (unsigned)r[p] : 0u;
    }
  }
  for (unsigned i = threadIdx.x; i < (unsigned)nq * (unsigned)(sizeof(QDesc) / 4); i += BLOCK)
    ((unsigned*)qd)[i] = ((const unsigned*)A->qd)[i];
  {
    const unsigned nb = (unsigned)A->seg0[kMaxCols];
    for (unsigned i = threadIdx.x; i < nb; i += BLOCK) lsegb[i] = A->segb[i];
    for (unsigned i = threadIdx.x; i < nb + (unsigned)kMaxCols; i += BLOCK) lsegm[i] = A->segm[i];
  }
  if (LMAP) {
    for (int s = 0; s < nslot; ++s) {
      const Slot& S = A->slot[s];
      if (S.lmap >= 0) {
        // entries 0..range (entry range holds P), packed little-end first
        const unsigned lg = (unsigned)S.lmap_log2, per = 32u >> lg;
        const unsigned words = (S.range + per) / per;
        for (unsigned w = threadIdx.x; w < words; w += BLOCK) {
          unsigned v = 0u;
          for (unsigned j = 0; j < per; ++j) {
            const unsigned k = w * per + j;
            const unsigned e = k > S.range ? 0u
                               : S.narrow  ? (unsigned)((const uint8_t*)S.pkey)[k]
                                           : (unsigned)((const uint16_t*)S.pkey)[k];
            v |= e << (j << lg);
          }
          lmaps[S.lmap + w] = v;
        }
      }
      if (S.lbits >= 0) {
        const unsigned words = S.range / 32u + 1u, have = 2u * ((S.range + 63u) / 64u);
        for (unsigned w = threadIdx.x; w < words; w += BLOCK) lmaps[S.lbits + w] = w < have ? S.bits[w] : 0u;
      }
    }
  }
  for (unsigned i = threadIdx.x; i < (unsigned)kSlots; i += BLOCK) hsum[i] = 0ull;
  for (unsigned i = threadIdx.x; i < kKeyWords; i += BLOCK) hkey[i] = 0u;
  if (threadIdx.x == 0) used = 0u;
  // a scalar query yields its one row even when no row passes
  if (blockIdx.x == 0 && (int)threadIdx.x < nq && A->qd[threadIdx.x].K == 0)
    atomicOr(gflag + A->qd[threadIdx.x].flag_off, 1u);
  __syncthreads();

  auto flush = [&]() {
    for (unsigned i = threadIdx.x; i < (unsigned)kSlots; i += BLOCK) {
      const unsigned k = hkey[i];
      if (k) {
        const unsigned key = k - 1u, q = key >> 22, gid = key & ((1u << 22) - 1u);
        atomicAdd(gsum + qd[q].sum_off + gid, hsum[i]);
        atomicOr(gflag + qd[q].flag_off + (gid >> 5), 1u << (gid & 31u));
        hkey[i] = 0u;
        hsum[i] = 0ull;
      }
    }
  };

  // the first cnt rows of the wave's queue, one per lane
  auto queue_flush = [&](unsigned cnt) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const unsigned l = threadIdx.x & 63u;
    if (l < cnt) {
      unsigned long long mask = wq[l];
      if (sizeof(M) == 8) mask |= (unsigned long long)wq[64u + l] << 32;
      const unsigned cw[3] = {wq[128u + l], wq[192u + l], wq[256u + l]};
      int mv[kMaxMetric];
#pragma unroll
      for (int k = 0; k < kMaxMetric; ++k) mv[k] = (int)wq[(5u + (unsigned)k) * 64u + l];
      for (M u = (M)mask; u; u &= u - (M)1) {
        const unsigned q = (unsigned)first_bit(u);
        const QDesc& d = qd[q];
        unsigned gid = 0u;
        for (int k = 0; k < d.K; ++k) {
          const unsigned kw = d.key[k], sl = (kw >> 13) & 7u;
          const unsigned w = sl < 2u ? cw[0] : sl < 4u ? cw[1] : cw[2];
          const unsigned c = (w >> ((sl & 1u) << 4)) & 0xFFFFu;
          gid += ((praw[(kw & 8191u) + c] >> ((kw >> 16) & 31u)) & ((1u << ((kw >> 21) & 31u)) - 1u)) * d.stride[k];
        }
        const unsigned long long m = (unsigned long long)(long long)sel_i(mv, d.metric, nmetric);
        const unsigned lo = d.dense;
        if (lo != kNoDense) {
          const unsigned i = lo + gid;
          atomicAdd(&hsum[i], m);
          atomicOr(&hkey[i >> 5], 1u << (i & 31u));
        } else {
          atomicAdd(gsum + d.sum_off + gid, m);
          atomicOr(gflag + d.flag_off + (gid >> 5), 1u << (gid & 31u));
        }
      }
    }
    // every lane has read its entry before the queue is written again
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  for (;;) {
    if (threadIdx.x < 64) {
      unsigned t0 = 0u;
      if (threadIdx.x == 0) t0 = atomicAdd(A->counter, 1u);
      t0 = __shfl(t0, 0);
      const long long lo = A->row_lo + (long long)t0 * A->tile_rows;
      const long long hi = min(lo + A->tile_rows, A->row_hi);
      bool full = false, part = false;
      unsigned cm = 0u;
      if (t0 < A->ntiles && (int)threadIdx.x < nq) {
        const long long r0 = A->qrow[2 * threadIdx.x], r1 = A->qrow[2 * threadIdx.x + 1];
        full = r0 <= lo && r1 >= hi;
        part = !full && r0 < hi && r1 > lo;
        if (full || part) cm = A->qcols[threadIdx.x];
      }
      const unsigned long long fm = __ballot(full), pm = __ballot(part);
      for (int d = 32; d > 0; d >>= 1) cm |= __shfl_xor(cm, d);
      if (threadIdx.x == 0) {
        cur = t0;
        tmask[0] = fm;
        tmask[1] = pm;
        tcols = cm;  // columns the tile's queries read
      }
    }
    __syncthreads();
    const unsigned t = cur;
    if (t >= A->ntiles) break;
    const M full = (M)tmask[0], part = (M)tmask[1], active = full | part;
    const unsigned colmask = tcols;
    // columns loaded for every group (all of them without staged loads)
    const unsigned lmask = ST ? colmask & A->tier0 : colmask;
    const long long lo = A->row_lo + (long long)t * A->tile_rows;  // a multiple of 4
    const long long hi = min(lo + A->tile_rows, A->row_hi);
    if (active) {
      const long long vend = (hi + 3) >> 2;
      // The loop runs while any lane of the wave has a group (lane 0 has the
      // wave's first), so every lane stays active for the wave-wide steps; a lane
      // past the tile loads nothing and passes no row (left <= 0 below).
      const long long ln = (long long)(threadIdx.x & 63u);
      const long long vw0 = (lo >> 2) + (long long)(threadIdx.x & ~63u);
      i32x4 x[kMaxCols];
      load_or_zero(A, lmask, vw0 + ln, vend, x);
      for (long long vw = vw0; vw < vend; vw += BLOCK) {
        const long long v = vw + ln;
        // PF: the next group's loads are issued before this group's lookups
        i32x4 xn[kMaxCols];
        if (PF) load_or_zero(A, colmask, v + BLOCK, vend, xn);
        else {
#pragma unroll
          for (int c = 0; c < kMaxCols; ++c) xn[c] = x[c];
        }
        const long long left = hi - v * 4;  // rows of this group inside the tile (<= 0 past it)
        M pass[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          M m = e < left ? full : (M)0;
          if (e < left) {
            const long long i = v * 4 + e;
            for (M pb = part; pb; pb &= pb - (M)1) {
              const int q = first_bit(pb);
              if (i >= A->qrow[2 * q] && i < A->qrow[2 * q + 1]) m |= (M)1 << q;
            }
          }
          pass[e] = m;
        }
        // fact predicates of the tile's queries, grouped by column: tested one by one
        // when the tile has few on the column, else through the column's value
        // segments (a search over its sorted bounds, then that segment's query mask)
        auto apply_preds = [&]() {
#pragma unroll
        for (int c = 0; c < kMaxCols; ++c) {
          const M cq = (M)A->colq[c] & active;
          if (!cq) continue;
          const int steps = A->segsteps[c];
          if (__popcll((unsigned long long)cq) <= steps + 1) {
            const int p1 = A->pred0[c + 1];
            for (int p = A->pred0[c]; p < p1; ++p) {
              const int q = A->pred[p].q;
              if (!((active >> q) & (M)1)) continue;
              const int plo = A->pred[p].lo, phi = A->pred[p].hi;
              const M keep = ~((M)1 << q);
#pragma unroll
              for (int e = 0; e < 4; ++e)
                if (x[c][e] < plo || x[c][e] > phi) pass[e] &= keep;
            }
          } else {
            const int b0 = A->seg0[c], nb = A->seg0[c + 1] - b0;
            const int* bounds = lsegb + b0;
            const unsigned long long* masks = lsegm + b0 + c;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int v = x[c][e];
              int pos = 0;  // number of bounds <= v
              for (int step = steps > 0 ? 1 << (steps - 1) : 0; step > 0; step >>= 1)
                if (pos + step <= nb && bounds[pos + step - 1] <= v) pos += step;
              pass[e] &= (M)masks[pos];
            }
          }
        }
        };
        // gated lookups of slots [s0, s1): each slot only for the rows some joining
        // query still wants
        unsigned cid[kMaxSlot][2];  // two rows' 16-bit profile ids per register
#pragma unroll
        for (int s = 0; s < kMaxSlot; ++s) cid[s][0] = cid[s][1] = 0u;
        auto gated_slots = [&](int s0, int s1) {
#pragma unroll
          for (int s = 0; s < kMaxSlot; ++s) {
            if (s < s0 || s >= s1 || s >= nslot || !((M)A->slot[s].joined & active)) continue;
            unsigned cs[4] = {0u, 0u, 0u, 0u};
            slot_lookup<M, LMAP>(A->slot[s], x, pass, use_bits, lmaps, cs);
            slot_apply<M>(A->slot[s], pmask, cs, pass);
            cid[s][0] = cs[0] | cs[1] << 16;
            cid[s][1] = cs[2] | cs[3] << 16;
          }
        };
        if (ST) {
          const int n0 = A->nslot0;
          gated_slots(0, n0);
          // the other columns, for lanes with a row some query still wants
          const bool any0 = (pass[0] | pass[1] | pass[2] | pass[3]) != (M)0;
          const unsigned rest = colmask & ~lmask;
#pragma unroll
          for (int c = 0; c < kMaxCols; ++c)
            if (rest & (1u << c)) x[c] = any0 ? __builtin_nontemporal_load((const i32x4*)A->cols[c] + v) : i32x4{0, 0, 0, 0};
          apply_preds();
          gated_slots(n0, kMaxSlot);
        } else {
        apply_preds();
        // joins: one profile lookup per slot for the rows some joining query still
        // wants.  PAR issues every slot's lookups before combining any (one memory
        // latency per row group); otherwise each slot is gated by the slots before
        // it (fewer lookups when the first joins are selective).
        if (PAR) {
          unsigned cs[kMaxSlot][4];
#pragma unroll
          for (int s = 0; s < kMaxSlot; ++s) {
#pragma unroll
            for (int e = 0; e < 4; ++e) cs[s][e] = 0u;
            if (s < nslot && ((M)A->slot[s].joined & active)) slot_lookup<M, LMAP>(A->slot[s], x, pass, use_bits, lmaps, cs[s]);
          }
#pragma unroll
          for (int s = 0; s < kMaxSlot; ++s) {
            if (s < nslot && ((M)A->slot[s].joined & active)) slot_apply<M>(A->slot[s], pmask, cs[s], pass);
            cid[s][0] = cs[s][0] | cs[s][1] << 16;
            cid[s][1] = cs[s][2] | cs[s][3] << 16;
          }
        } else {
          gated_slots(0, kMaxSlot);
        }
        }  // not staged
        const M lane_any = pass[0] | pass[1] | pass[2] | pass[3];
        bool agg_here = lane_any != (M)0;
        if (Q) {
          unsigned long long qb[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) qb[e] = __ballot(pass[e] != (M)0);
          const unsigned n = (unsigned)(__popcll(qb[0]) + __popcll(qb[1]) + __popcll(qb[2]) + __popcll(qb[3]));
          agg_here = n >= kQueueInline;  // a dense group: aggregated in place
          if (n != 0u && !agg_here) {
            if (qn + n > 64u) {
              queue_flush(qn);
              qn = 0u;
            }
            unsigned base = qn;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              if (pass[e] != (M)0) {
                const unsigned pos =
                    base + __builtin_amdgcn_mbcnt_hi((unsigned)(qb[e] >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)qb[e], 0u));
                wq[pos] = (unsigned)pass[e];
                if (sizeof(M) == 8) wq[64u + pos] = (unsigned)((unsigned long long)pass[e] >> 32);
                unsigned c[kMaxSlot];
#pragma unroll
                for (int sl = 0; sl < kMaxSlot; ++sl) c[sl] = (cid[sl][e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
                wq[128u + pos] = c[0] | c[1] << 16;
                wq[192u + pos] = c[2] | c[3] << 16;
                wq[256u + pos] = c[4] | c[5] << 16;
#pragma unroll
                for (int k = 0; k < kMaxMetric; ++k) {
                  if (k < nmetric) {
                    const int op = A->mop[k];
                    const unsigned a = (unsigned)pick(x, A->ma[k], e), b = (unsigned)pick(x, A->mb[k], e);
                    wq[(5u + (unsigned)k) * 64u + pos] = op == MOP_COL ? a : op == MOP_SUB ? a - b : op == MOP_ADD ? a + b : a * b;
                  }
                }
              }
              base += (unsigned)__popcll(qb[e]);
            }
            qn += n;
          }
        }
        if (agg_here) {
        // distinct metric expressions
        int mval[kMaxMetric][4];
#pragma unroll
        for (int k = 0; k < kMaxMetric; ++k) {
#pragma unroll
          for (int e = 0; e < 4; ++e) mval[k][e] = 0;
          if (k < nmetric) {
            const int op = A->mop[k], ca = A->ma[k], cb = A->mb[k];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const unsigned a = (unsigned)pick(x, ca, e), b = (unsigned)pick(x, cb, e);
              mval[k][e] = op == MOP_COL ? (int)a : op == MOP_SUB ? (int)(a - b) : op == MOP_ADD ? (int)(a + b)
                                                                                    : (int)(a * b);
            }
          }
        }
        if (DENSE) {
          unsigned c01[kMaxSlot], c23[kMaxSlot];
#pragma unroll
          for (int s = 0; s < kMaxSlot; ++s) {
            c01[s] = cid[s][0];
            c23[s] = cid[s][1];
          }
          for (M u = pass[0] | pass[1] | pass[2] | pass[3]; u; u &= u - (M)1) {
            const unsigned q = (unsigned)first_bit(u);
            const QDesc& d = qd[q];
            unsigned gid[4] = {0u, 0u, 0u, 0u};
            for (int k = 0; k < d.K; ++k) {
              const unsigned kw = d.key[k], st = d.stride[k];
              const unsigned off = kw & 8191u, sh = (kw >> 16) & 31u, msk = (1u << ((kw >> 21) & 31u)) - 1u;
              const unsigned w01 = sel<kMaxSlot>(c01, (kw >> 13) & 7u, nslot);
              const unsigned w23 = sel<kMaxSlot>(c23, (kw >> 13) & 7u, nslot);
              gid[0] += ((praw[off + (w01 & 0xFFFFu)] >> sh) & msk) * st;
              gid[1] += ((praw[off + (w01 >> 16)] >> sh) & msk) * st;
              gid[2] += ((praw[off + (w23 & 0xFFFFu)] >> sh) & msk) * st;
              gid[3] += ((praw[off + (w23 >> 16)] >> sh) & msk) * st;
            }
            bool has[4];
            long long mv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              int ms[kMaxMetric];
#pragma unroll
              for (int k = 0; k < kMaxMetric; ++k) ms[k] = mval[k][e];
              has[e] = ((pass[e] >> q) & (M)1) != (M)0;
              mv[e] = (long long)sel_i(ms, d.metric, nmetric);
            }
            // the lane's rows in one group of the query are added up first
#pragma unroll
            for (int e = 1; e < 4; ++e)
#pragma unroll
              for (int f = 0; f < e; ++f)
                if (has[e] && has[f] && gid[e] == gid[f]) {
                  mv[f] += mv[e];
                  has[e] = false;
                }
            const unsigned lo = d.dense;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              if (!has[e]) continue;
              if (lo != kNoDense) {
                const unsigned i = lo + gid[e];
                atomicAdd(&hsum[i], (unsigned long long)mv[e]);
                atomicOr(&hkey[i >> 5], 1u << (i & 31u));
              } else {
                atomicAdd(gsum + d.sum_off + gid[e], (unsigned long long)mv[e]);
                atomicOr(gflag + d.flag_off + (gid[e] >> 5), 1u << (gid[e] & 31u));
              }
            }
          }
        } else {
        // this lane's (row, query) pairs
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          unsigned cs[kMaxSlot];
#pragma unroll
          for (int s = 0; s < kMaxSlot; ++s) cs[s] = (cid[s][e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
          int ms[kMaxMetric];
#pragma unroll
          for (int k = 0; k < kMaxMetric; ++k) ms[k] = mval[k][e];
          for (M bits = pass[e]; bits; bits &= bits - (M)1) {
            const unsigned q = (unsigned)first_bit(bits);
            const QDesc& d = qd[q];
            unsigned gid = 0u;
            for (int k = 0; k < d.K; ++k) {
              const unsigned kw = d.key[k];
              const unsigned c = sel<kMaxSlot>(cs, (kw >> 13) & 7u, nslot);
              const unsigned raw = praw[(kw & 8191u) + c];
              const unsigned nb = (kw >> 21) & 31u;
              gid += ((raw >> ((kw >> 16) & 31u)) & ((1u << nb) - 1u)) * d.stride[k];
            }
            const unsigned long long mv = (unsigned long long)(long long)sel_i(ms, d.metric, nmetric);
            const unsigned key = ((q << 22) | gid) + 1u;
            unsigned h = (key * 2654435761u) >> (32 - HBITS);
            bool done = false;
            for (int probe = 0; probe < 8; ++probe) {
              unsigned k = hkey[h];
              if (k == 0u) {
                const unsigned old = atomicCAS(&hkey[h], 0u, key);
                if (old == 0u) atomicAdd(&used, 1u);
                k = old == 0u ? key : old;
              }
              if (k == key) {
                atomicAdd(&hsum[h], mv);
                done = true;
                break;
              }
              h = (h + 1u) & (unsigned)(kSlots - 1);
            }
            if (!done) {
              atomicAdd(gsum + d.sum_off + gid, mv);
              atomicOr(gflag + d.flag_off + (gid >> 5), 1u << (gid & 31u));
            }
          }
        }
        }  // hash aggregation
        }  // some row of the lane passes
        // the next group: already in flight (PF) or requested now
#pragma unroll
        for (int c = 0; c < kMaxCols; ++c) x[c] = xn[c];
        if (!PF) load_or_zero(A, lmask, v + BLOCK, vend, x);
      }
    }
    __syncthreads();
    if (!DENSE && used > (unsigned)kSlots / 2u) {
      flush();
      __syncthreads();
      if (threadIdx.x == 0) used = 0u;
    }
  }
  if (Q && qn) queue_flush(qn);
  __syncthreads();
  if (!DENSE) {
    flush();
    return;
  }
  // dense tables: every group with its existence bit set goes to the HBM tables
  for (int q = 0; q < nq; ++q) {
    const unsigned lo = qd[q].dense;
    if (lo == kNoDense) continue;
    const unsigned G = qd[q].G;
    for (unsigned g = threadIdx.x; g < G; g += BLOCK) {
      const unsigned i = lo + g;
      if ((hkey[i >> 5] >> (i & 31u)) & 1u) {
        atomicAdd(gsum + qd[q].sum_off + g, hsum[i]);
        atomicOr(gflag + qd[q].flag_off + (g >> 5), 1u << (g & 31u));
      }
    }
  }
}

// ------------------------------------------------------------------ emission
// Block b of the flattened emission grid belongs to the query whose block range holds it.
__device__ __forceinline__ int owner(const Emit* e, int nq, unsigned long long b) {
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (e[mid].block0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Zeroes the pass's group tables and its tile counter.
__global__ void k_pass_clear(const ScanArgs* __restrict__ A) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long* sums = A->sums;
  unsigned* flags = A->flags;
  for (uint64_t i = t0; i < A->sum_words; i += stride) sums[i] = 0ull;
  for (uint64_t i = t0; i < A->flag_words; i += stride) flags[i] = 0u;
  // the profile masks k_profile_masks ORs into
  for (int s = 0; s < A->nslot; ++s) {
    unsigned long long* m = (unsigned long long*)A->slot[s].mask;
    for (uint64_t i = t0; i < A->slot[s].P; i += stride) m[i] = 0ull;
  }
  if (t0 == 0) {
    *A->counter = 0u;
    if (A->dropq) *A->dropq = 0ull;
  }
}

// Queries (of `joined`) that keep the dimension row of entry e of a checked slot.
__device__ __forceinline__ unsigned long long check_member(const DropCheck& D, unsigned e) {
  const unsigned off = (unsigned)(D.key[e] - D.kmin);
  if (off >= D.range) return 0ull;
  const unsigned p = D.narrow ? ((const unsigned char*)D.pkey)[off] : ((const unsigned short*)D.pkey)[off];
  return p < D.P ? D.mask[p] & D.joined : 0ull;
}

__global__ void k_drop_check(const ScanArgs* __restrict__ A) {
  const int nchk = A->nchk;
  const unsigned stride = gridDim.x * blockDim.x;
  for (int k = 0; k < nchk; ++k) {
    const DropCheck& D = A->chk[k];
    for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < D.clusters; c += stride) {
      const unsigned e0 = D.start[c], e1 = D.start[c + 1];
      // members per query as bit-sliced 4-bit counters (c4: 16 or more)
      unsigned long long c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;
      unsigned base = 0xFFFFFFFFu;
      for (unsigned e = e0; e < e1; ++e) {
        base = min(base, D.home[e]);
        unsigned long long carry = check_member(D, e), t;
        t = c0 & carry; c0 ^= carry; carry = t;
        t = c1 & carry; c1 ^= carry; carry = t;
        t = c2 & carry; c2 ^= carry; carry = t;
        t = c3 & carry; c3 ^= carry; carry = t;
        c4 |= carry;
      }
      unsigned long long cand = c4 | (c3 & (c2 | (c1 & c0)));  // 11 or more
      while (cand) {
        const int q = __ffsll((long long)cand) - 1;
        cand &= cand - 1ull;
        // the query's inserts in row order; the cluster spans at most kCheckSpan
        // slots from `base` (checked when the slot is set up), so no position wraps
        unsigned long long occ[kCheckSpan / 64] = {};
        for (unsigned e = e0; e < e1; ++e) {
          if (!((check_member(D, e) >> q) & 1ull)) continue;
          unsigned pos = D.home[e] - base;
          bool placed = false;
          for (int t = 0; t < 10; ++t, ++pos) {
            const unsigned long long bit = 1ull << (pos & 63u);
            if (!(occ[pos >> 6] & bit)) {
              occ[pos >> 6] |= bit;
              placed = true;
              break;
            }
          }
          if (!placed) {
            atomicOr(A->dropq, 1ull << q);
            break;
          }
        }
      }
    }
  }
}

__global__ void k_emit_count(const Emit* __restrict__ e, const ScanArgs* __restrict__ A,
                             unsigned long long* counts) {
  __shared__ unsigned part[kBlock / 64];
  const int nq = A->nq;
  const unsigned long long total_blocks = A->total_blocks;
  for (unsigned long long b = blockIdx.x; b < total_blocks; b += gridDim.x) {
    int q = owner(e, nq, b);
    unsigned long long w = (b - e[q].block0) * kBlock + threadIdx.x;
    unsigned long long nw = (e[q].G + 31u) / 32u;
    unsigned c = w < nw ? (unsigned)__popc(e[q].flags[w]) : 0u;
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int k = 0; k < kBlock / 64; ++k) t += part[k];
      counts[b] = t;
    }
    __syncthreads();
  }
}

// One workgroup per query: exclusive scan of its block counts, global row offsets.
__global__ void k_emit_scan(const Emit* __restrict__ e, const ScanArgs* __restrict__ A,
                            const unsigned long long* counts, unsigned long long* offsets,
                            unsigned long long* totals) {
  const int q = blockIdx.x;
  if (q >= A->nq) return;
  __shared__ unsigned long long buf[kBlock];
  unsigned long long carry = 0;
  for (unsigned long long base = 0; base < e[q].nblocks; base += kBlock) {
    unsigned long long i = base + threadIdx.x;
    unsigned long long v = i < e[q].nblocks ? counts[e[q].block0 + i] : 0ull;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
      unsigned long long t = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0ull;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < e[q].nblocks) offsets[e[q].block0 + i] = carry + buf[threadIdx.x] - v;
    unsigned long long blk = buf[kBlock - 1];
    __syncthreads();
    carry += blk;
  }
  if (threadIdx.x == 0) totals[q] = carry;
}

__global__ void k_emit_rows(const Emit* __restrict__ e, int nq, unsigned long long total_blocks,
                            const unsigned long long* __restrict__ offsets, int32_t* keys, int64_t* sums,
                            unsigned long long arena_rows) {
  __shared__ unsigned pre[kBlock];
  for (unsigned long long b = blockIdx.x; b < total_blocks; b += gridDim.x) {
    int q = owner(e, nq, b);
    const Emit& E = e[q];
    unsigned long long w = (b - E.block0) * kBlock + threadIdx.x;
    unsigned long long nw = (E.G + 31u) / 32u;
    unsigned bits = w < nw ? E.flags[w] : 0u;
    unsigned c = (unsigned)__popc(bits);
    pre[threadIdx.x] = c;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
      unsigned t = threadIdx.x >= (unsigned)d ? pre[threadIdx.x - d] : 0u;
      __syncthreads();
      pre[threadIdx.x] += t;
      __syncthreads();
    }
    unsigned long long row = E.out0 + offsets[b] + pre[threadIdx.x] - c;
    while (bits) {
      int bit = __ffs(bits) - 1;
      bits &= bits - 1;
      unsigned long long g = w * 32 + (unsigned long long)bit;
      for (int k = 0; k < E.K; ++k) {
        unsigned code = (unsigned)((g / E.stride[k]) % E.n[k]);
        keys[(unsigned long long)k * arena_rows + row] = E.dict[k][code];
      }
      sums[row] = (int64_t)E.sums[g];
      row++;
    }
    __syncthreads();
  }
}

inline unsigned grid_rows(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g == 0) g = 1;
  return (unsigned)(g > 4096 ? 4096 : g);
}

}  // namespace

void launch_profile_raw(const RawArgs& a, uint64_t n, int32_t* raw, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_profile_raw, dim3(grid_rows(n)), dim3(kBlock), 0, s, a, n, raw);
  hip_check(hipGetLastError(), "k_profile_raw");
}

void launch_profile_scatter(const int32_t* keys, uint64_t n, int32_t kmin, const int32_t* pcode, void* pkey,
                            int narrow, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_profile_scatter, dim3(grid_rows(n)), dim3(kBlock), 0, s, keys, n, kmin, pcode, pkey, narrow);
  hip_check(hipGetLastError(), "k_profile_scatter");
}

void launch_string_exact(const char* col, int size, uint64_t n, const int32_t* codes, const int32_t* rep,
                         unsigned* flag, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_string_exact, dim3(grid_rows(n)), dim3(kBlock), 0, s, col, size, n, codes, rep, flag);
  hip_check(hipGetLastError(), "k_string_exact");
}

// Every launch below except the emission of rows has a grid and arguments that do
// not depend on the batch (sizes are read from the device descriptors), so the
// sequence clear -> masks -> scan -> emission count can be captured once in a
// hipGraph and replayed for any batch whose descriptors are uploaded in place.
constexpr unsigned kClearGrid = 1024;
constexpr unsigned kEmitGrid = 2048;

void launch_pass_clear(const ScanArgs* d_args, hipStream_t s) {
  hipLaunchKernelGGL(k_pass_clear, dim3(kClearGrid), dim3(kBlock), 0, s, d_args);
  hip_check(hipGetLastError(), "k_pass_clear");
}

void launch_profile_masks(const MaskArgs* d_margs, hipStream_t s) {
  hipLaunchKernelGGL(k_profile_masks, dim3(kMaxProfiles / kBlock, kMaxSlot, kMaxTerms / kTermsPerThread), dim3(kBlock),
                     0, s, d_margs);
  hip_check(hipGetLastError(), "k_profile_masks");
}

void launch_drop_check(const ScanArgs* d_args, hipStream_t s) {
  hipLaunchKernelGGL(k_drop_check, dim3(kClearGrid), dim3(kBlock), 0, s, d_args);
  hip_check(hipGetLastError(), "k_drop_check");
}

void launch_slot_bits(const ScanArgs* d_args, hipStream_t s) {
  hipLaunchKernelGGL(k_slot_bits, dim3(kClearGrid), dim3(kBlock), 0, s, d_args);
  hip_check(hipGetLastError(), "k_slot_bits");
}

namespace {
struct Variant {
  int block, hbits;
  bool par, pf, mask32;
  int agg;
  bool lmap, queue, staged;
  const void* fn;
};
#define SVM_SCAN_VARIANT(B, H, PAR, PF, M32, D, L, Q, ST)                                                            \
  {                                                                                                                \
    B, H, PAR, PF, M32, D, L, Q, ST,                                                                               \
        (const void*)k_shared_scan<B, H, PAR, PF, std::conditional_t<M32, unsigned, unsigned long long>, D, L, Q, ST> \
  }
// (threads, log2 LDS accumulator slots, parallel lookups, prefetch, 32-bit masks,
//  aggregation, LDS maps, queue, staged loads)
const Variant kVariants[] = {
    SVM_SCAN_VARIANT(512, 11, false, false, false, 0, false, false, false),
    SVM_SCAN_VARIANT(512, 11, false, false, true, 0, false, false, false),
    SVM_SCAN_VARIANT(512, 11, true, false, false, 0, false, false, false),
    SVM_SCAN_VARIANT(512, 11, true, false, true, 0, false, false, false),
    SVM_SCAN_VARIANT(512, 11, false, true, false, 0, false, false, false),
    SVM_SCAN_VARIANT(512, 11, false, true, true, 0, false, false, false),
    SVM_SCAN_VARIANT(256, 10, false, false, false, 0, false, false, false),
    SVM_SCAN_VARIANT(256, 10, false, false, true, 0, false, false, false),
    SVM_SCAN_VARIANT(1024, 12, false, false, false, 1, true, false, false),
    SVM_SCAN_VARIANT(1024, 12, false, false, true, 1, true, false, false),
    SVM_SCAN_VARIANT(1024, 12, false, false, false, 1, true, true, false),
    SVM_SCAN_VARIANT(1024, 12, false, false, true, 1, true, true, false),
    SVM_SCAN_VARIANT(1024, 12, false, false, false, 1, true, true, true),
    SVM_SCAN_VARIANT(1024, 12, false, false, true, 1, true, true, true),
    SVM_SCAN_VARIANT(512, 11, false, false, false, 0, false, false, true),
    SVM_SCAN_VARIANT(512, 11, false, false, true, 0, false, false, true),
};
#undef SVM_SCAN_VARIANT
// dynamic LDS of the variants with LDS maps: the CU's 160 KB less a margin for the
// kernel's static variables
constexpr size_t kLdsMapsTotal = 160 * 1024 - 512;
}  // namespace

int scan_variants() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }
int scan_variant_block(int v) { return kVariants[v].block; }
int scan_variant_hash_slots(int v) { return 1 << kVariants[v].hbits; }
bool scan_variant_parallel(int v) { return kVariants[v].par; }
bool scan_variant_prefetch(int v) { return kVariants[v].pf; }
bool scan_variant_mask32(int v) { return kVariants[v].mask32; }
bool scan_variant_dense(int v) { return kVariants[v].agg != 0; }
bool scan_variant_staged(int v) { return kVariants[v].staged; }

void assign_dense(QDesc* qd, int nq, unsigned slots) {
  // queries with the fewest groups first, while their tables fit
  int order[kMaxBatch];
  for (int q = 0; q < nq; ++q) order[q] = q;
  std::stable_sort(order, order + nq, [&](int a, int b) { return qd[a].G < qd[b].G; });
  unsigned used = 0;
  for (int k = 0; k < nq; ++k) {
    QDesc& d = qd[order[k]];
    if (d.G <= slots - used) {
      d.dense = used;
      used += d.G;
    } else {
      d.dense = kNoDense;
    }
  }
}

// LDS of a variant before its map area
size_t fixed_lds(int variant, unsigned nprof_cap) {
  const size_t slots = (size_t)1 << kVariants[variant].hbits;
  const size_t key_words = kVariants[variant].agg ? slots / 32 : slots;
  const size_t queue = kVariants[variant].queue ? (size_t)kVariants[variant].block * kQWords * 4 : 0;
  return slots * 8 + key_words * 4 + (size_t)nprof_cap * 12 + (size_t)kMaxBatch * sizeof(QDesc) +
         (size_t)(kMaxSegBounds + kMaxCols) * 8 + (size_t)kMaxSegBounds * 4 + queue;
}

size_t shared_scan_lds(int variant, unsigned nprof_cap) {
  return kVariants[variant].lmap ? kLdsMapsTotal : fixed_lds(variant, nprof_cap);
}

bool scan_variant_lds_maps(int v) { return kVariants[v].lmap; }

void assign_lds_maps(ScanArgs* A, int variant, unsigned nprof_cap) {
  for (int s = 0; s < A->nslot; ++s) {
    A->slot[s].lmap = -1;
    A->slot[s].lmap_log2 = 0;
    A->slot[s].lbits = -1;
  }
  if (!kVariants[variant].lmap) return;
  const size_t fixed = fixed_lds(variant, nprof_cap);
  size_t left = kLdsMapsTotal > fixed ? (kLdsMapsTotal - fixed) / 4 : 0;  // words
  size_t used = 0;
  int order[kMaxSlot];
  size_t map_words[kMaxSlot];
  int lg[kMaxSlot];
  for (int s = 0; s < A->nslot; ++s) {
    order[s] = s;
    const uint64_t entries = (uint64_t)A->slot[s].range + 1;  // entry range holds P
    lg[s] = A->slot[s].P + 1 <= 16 ? 2 : A->slot[s].P + 1 <= 256 ? 3 : 4;
    const uint64_t per = 32u >> lg[s];
    map_words[s] = (size_t)((entries + per - 1) / per);
  }
  std::stable_sort(order, order + A->nslot, [&](int a, int b) { return map_words[a] < map_words[b]; });
  for (int k = 0; k < A->nslot; ++k) {
    Slot& S = A->slot[order[k]];
    if (map_words[order[k]] <= left - used) {
      S.lmap = (int)used;
      S.lmap_log2 = lg[order[k]];
      used += map_words[order[k]];
    }
  }
  for (int k = 0; k < A->nslot; ++k) {
    Slot& S = A->slot[order[k]];
    const size_t words = (size_t)S.range / 32 + 1;
    if (S.lmap < 0 && words <= left - used) {
      S.lbits = (int)used;
      used += words;
    }
  }
}

void prepare_shared_scan() {
  static const bool done = [] {
    if (kDefaultVariantMask32 >= scan_variants() || !kVariants[kDefaultVariantMask32].mask32)
      fail(SVM_ERR_STATE, "the default shared-scan variant for 32-bit masks has 64-bit masks");
    for (int v = 0; v < scan_variants(); ++v)
      hip_check(hipFuncSetAttribute(kVariants[v].fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)shared_scan_lds(v, kMaxProfiles)),
                "k_shared_scan LDS limit");
    return true;
  }();
  (void)done;
}

void launch_shared_scan(const ScanArgs* d_args, int variant, unsigned nprof_cap, int grid, hipStream_t s) {
  const size_t lds = shared_scan_lds(variant, nprof_cap);
  const ScanArgs* arg = d_args;
  void* params[] = {&arg};
  hip_check(hipLaunchKernel(kVariants[variant].fn, dim3(grid), dim3(kVariants[variant].block), params, lds, s),
            "k_shared_scan");
}

void launch_emit_count(const Emit* d_e, const ScanArgs* d_args, unsigned long long* d_counts,
                       unsigned long long* d_offsets, unsigned long long* d_totals, hipStream_t s) {
  hipLaunchKernelGGL(k_emit_count, dim3(kEmitGrid), dim3(kBlock), 0, s, d_e, d_args, d_counts);
  hipLaunchKernelGGL(k_emit_scan, dim3(kMaxBatch), dim3(kBlock), 0, s, d_e, d_args, d_counts, d_offsets, d_totals);
  hip_check(hipGetLastError(), "k_emit_count");
}

void launch_emit_rows(const Emit* d_e, int nq, unsigned long long total_blocks,
                      const unsigned long long* d_offsets, int32_t* keys_arena, int64_t* sums_arena,
                      unsigned long long arena_rows, hipStream_t s) {
  if (nq <= 0 || !arena_rows) return;
  unsigned g = (unsigned)(total_blocks < 4096 ? (total_blocks ? total_blocks : 1) : 4096);
  hipLaunchKernelGGL(k_emit_rows, dim3(g), dim3(kBlock), 0, s, d_e, nq, total_blocks, d_offsets, keys_arena,
                     sums_arena, arena_rows);
  hip_check(hipGetLastError(), "k_emit_rows");
}

}  // namespace bk
}  // namespace svm
