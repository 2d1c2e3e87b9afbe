// batch_kernels.hip — parameter-driven kernels of the batched path: fused
// dimension filter + map build, the shared scan over (tile, query) work items, and
// batched ordered group emission.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "svm_internal.h"

namespace svm {
namespace bk {

namespace {

constexpr int kBlock = 256;
typedef int i32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int ref_strcmp(const char* a, const char* b, int n) {
  for (int k = 0; k < n; ++k) {
    char x = a[k], y = b[k];
    if (x == '\0' || y == '\0') return 0;
    if (x != y) return (int)(signed char)x - (int)(signed char)y;
  }
  return 0;
}

__device__ __forceinline__ bool dim_pass(const DimPred& p, unsigned long long r) {
  const char* v = p.col + r * (unsigned long long)p.size;
  if (p.kind == DP_RANGE) {
    int x = p.type == Int16 ? (int)*(const short*)v : p.type == Byte ? (int)*(const signed char*)v : *(const int*)v;
    return x >= p.lo && x <= p.hi;
  }
  int c = ref_strcmp(v, p.lit, p.n);
  return p.kind == DP_STREQ ? c == 0 : c != 0;
}

__global__ void k_map_build(const MapBuild* __restrict__ builds) {
  const MapBuild& b = builds[blockIdx.y];
  const int sh = b.wbits == 1 ? 5 : b.wbits == 2 ? 4 : b.wbits == 4 ? 3 : b.wbits == 8 ? 2 : b.wbits == 16 ? 1 : 0;
  for (unsigned long long r = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; r < b.n;
       r += (unsigned long long)gridDim.x * blockDim.x) {
    bool ok = true;
    for (int p = 0; p < b.npred; ++p) ok = ok && dim_pass(b.pred[p], r);
    if (!ok) continue;
    unsigned off = (unsigned)(b.keys[r] - b.kmin);
    if (off >= b.range) continue;
    unsigned val = b.codes ? (unsigned)b.codes[r] + 1u : 1u;
    unsigned per = 32u >> (5 - sh);  // entries per word (1 when wbits == 32)
    unsigned pos = sh == 0 ? 0u : (off & (per - 1u)) * (unsigned)b.wbits;
    atomicOr(&b.map[off >> sh], val << pos);
  }
}

__device__ __forceinline__ unsigned map_get(const unsigned* m, unsigned off, int w, int sh) {
  if (w == 32) return m[off];
  unsigned per = 32u >> (5 - sh);
  return (m[off >> sh] >> ((off & (per - 1u)) * (unsigned)w)) & ((1u << w) - 1u);
}

struct ColPtrs {
  const void* c[kMaxCols];
};

__device__ __forceinline__ int pick(const i32x4 (&x)[kMaxCols], int c, int e) {
  int v = x[0][e];
#pragma unroll
  for (int k = 1; k < kMaxCols; ++k) v = c == k ? x[k][e] : v;
  return v;
}

__global__ void __launch_bounds__(kBlock) k_shared_scan(const Query* __restrict__ qs, const Item* __restrict__ items,
                                                        unsigned nitems, unsigned* counter, ColPtrs cols,
                                                        long long tile_rows) {
  __shared__ unsigned long long lsum[kDenseG];
  __shared__ unsigned lflag[kDenseG / 32];
  __shared__ unsigned hkey[kHashSlots];
  __shared__ unsigned long long hsum[kHashSlots];
  __shared__ long long red[kBlock / 64];
  __shared__ unsigned cur;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(counter, 1u);
    __syncthreads();
    const unsigned it = cur;
    if (it >= nitems) break;
    const Item w = items[it];
    const Query& Q = qs[w.q];
    const long long lo = max((long long)w.tile * tile_rows, Q.row0);
    const long long hi = min(((long long)w.tile + 1) * tile_rows, Q.row1);
    const int agg = Q.agg, npred = Q.npred, njoin = Q.njoin, mop = Q.mop, ma = Q.ma, mb = Q.mb;
    const unsigned G = Q.G;
    unsigned mask = 1u << ma;
    if (mop != MOP_COL) mask |= 1u << mb;
    for (int p = 0; p < npred; ++p) mask |= 1u << Q.pcol[p];
    for (int j = 0; j < njoin; ++j) mask |= 1u << Q.jcol[j];
    int pcol[kMaxPred], plo[kMaxPred], phi[kMaxPred];
#pragma unroll
    for (int p = 0; p < kMaxPred; ++p) {
      pcol[p] = Q.pcol[p];
      plo[p] = Q.plo[p];
      phi[p] = Q.phi[p];
    }
    int jcol[kMaxJoin], jkmin[kMaxJoin], jw[kMaxJoin], jsh[kMaxJoin], jcode[kMaxJoin];
    unsigned jrange[kMaxJoin], jstride[kMaxJoin];
    const unsigned* jmap[kMaxJoin];
#pragma unroll
    for (int j = 0; j < kMaxJoin; ++j) {
      jcol[j] = Q.jcol[j];
      jkmin[j] = Q.jkmin[j];
      jw[j] = Q.jw[j];
      jsh[j] = jw[j] == 1 ? 5 : jw[j] == 2 ? 4 : jw[j] == 4 ? 3 : jw[j] == 8 ? 2 : jw[j] == 16 ? 1 : 0;
      jcode[j] = Q.jcode[j];
      jrange[j] = Q.jrange[j];
      jstride[j] = Q.jstride[j];
      jmap[j] = Q.jmap[j];
    }
    unsigned long long* gsum = Q.sums;
    unsigned* gflag = Q.flags;
    if (agg == BAGG_DENSE) {
      for (unsigned g = threadIdx.x; g < G; g += kBlock) lsum[g] = 0ull;
      for (unsigned g = threadIdx.x; g < (G + 31u) / 32u; g += kBlock) lflag[g] = 0u;
    } else if (agg == BAGG_HASH) {
      for (unsigned g = threadIdx.x; g < (unsigned)kHashSlots; g += kBlock) { hkey[g] = 0u; hsum[g] = 0ull; }
    }
    __syncthreads();
    long long acc = 0;
    for (long long v = (lo >> 2) + threadIdx.x; v < ((hi + 3) >> 2); v += kBlock) {
      i32x4 x[kMaxCols];
#pragma unroll
      for (int c = 0; c < kMaxCols; ++c) {
        if (mask & (1u << c)) x[c] = __builtin_nontemporal_load((const i32x4*)cols.c[c] + v);
        else x[c] = i32x4{0, 0, 0, 0};
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long long i = v * 4 + e;
        bool cond = i >= lo && i < hi;
        for (int p = 0; p < npred && cond; ++p) {
          int val = pick(x, pcol[p], e);
          cond = val >= plo[p] && val <= phi[p];
        }
        unsigned gid = 0;
        for (int j = 0; j < njoin && cond; ++j) {
          unsigned off = (unsigned)(pick(x, jcol[j], e) - jkmin[j]);
          unsigned pv = off < jrange[j] ? map_get(jmap[j], off, jw[j], jsh[j]) : 0u;
          cond = pv != 0u;
          if (jcode[j]) gid += (pv - 1u) * jstride[j];
        }
        if (!cond) continue;
        int a = pick(x, ma, e);
        int b = mop == MOP_COL ? 0 : pick(x, mb, e);
        int r = mop == MOP_COL ? a
              : mop == MOP_SUB ? (int)((unsigned)a - (unsigned)b)
              : mop == MOP_ADD ? (int)((unsigned)a + (unsigned)b)
                               : (int)((unsigned)a * (unsigned)b);
        long long m = (long long)r;
        if (agg == BAGG_SCALAR) {
          acc += m;
        } else if (agg == BAGG_DENSE) {
          atomicAdd(&lsum[gid], (unsigned long long)m);
          atomicOr(&lflag[gid >> 5], 1u << (gid & 31u));
        } else {
          unsigned h = (gid * 2654435761u) >> 22;  // 10 bits
          int slot = -1;
          for (int t = 0; t < 16; ++t) {
            unsigned k = hkey[h];
            if (k == gid + 1u) { slot = (int)h; break; }
            if (k == 0u) {
              unsigned old = atomicCAS(&hkey[h], 0u, gid + 1u);
              if (old == 0u || old == gid + 1u) { slot = (int)h; break; }
            }
            h = (h + 1u) & (kHashSlots - 1u);
          }
          if (slot >= 0) atomicAdd(&hsum[slot], (unsigned long long)m);
          else {
            atomicAdd(&gsum[gid], (unsigned long long)m);
            atomicOr(&gflag[gid >> 5], 1u << (gid & 31u));
          }
        }
      }
    }
    if (agg == BAGG_SCALAR) {
      for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (agg == BAGG_SCALAR) {
      if (threadIdx.x == 0) {
        long long t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += red[q];
        atomicAdd(&gsum[0], (unsigned long long)t);
        atomicOr(&gflag[0], 1u);
      }
    } else if (agg == BAGG_DENSE) {
      for (unsigned g = threadIdx.x; g < G; g += kBlock)
        if ((lflag[g >> 5] >> (g & 31u)) & 1u) atomicAdd(&gsum[g], lsum[g]);
      for (unsigned g = threadIdx.x; g < (G + 31u) / 32u; g += kBlock)
        if (lflag[g]) atomicOr(&gflag[g], lflag[g]);
    } else {
      for (unsigned s = threadIdx.x; s < (unsigned)kHashSlots; s += kBlock) {
        unsigned k = hkey[s];
        if (k) {
          atomicAdd(&gsum[k - 1u], hsum[s]);
          atomicOr(&gflag[(k - 1u) >> 5], 1u << ((k - 1u) & 31u));
        }
      }
    }
    __syncthreads();
  }
}

// Block b of the flattened emission grid belongs to the query whose block range holds it.
__device__ __forceinline__ int owner(const Emit* e, int nq, unsigned long long b) {
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (e[mid].block0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_emit_count(const Emit* __restrict__ e, int nq, unsigned long long total_blocks,
                             unsigned long long* counts) {
  __shared__ unsigned part[kBlock / 64];
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
__global__ void k_emit_scan(const Emit* __restrict__ e, int nq, const unsigned long long* counts,
                            unsigned long long* offsets, unsigned long long* totals) {
  const int q = blockIdx.x;
  if (q >= nq) return;
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

}  // namespace

void launch_map_build(const MapBuild* d_builds, int nbuild, unsigned long long max_rows, hipStream_t s) {
  if (nbuild <= 0) return;
  unsigned long long gx = (max_rows + kBlock - 1) / kBlock;
  if (gx > 1024) gx = 1024;
  if (gx == 0) gx = 1;
  hipLaunchKernelGGL(k_map_build, dim3((unsigned)gx, (unsigned)nbuild), dim3(kBlock), 0, s, d_builds);
  hip_check(hipGetLastError(), "k_map_build");
}

void launch_shared_scan(const Query* d_q, const Item* d_items, unsigned nitems, unsigned* d_counter,
                        const void* const* cols, long long tile_rows, int grid, hipStream_t s) {
  ColPtrs cp;
  for (int c = 0; c < kMaxCols; ++c) cp.c[c] = cols[c];
  hipLaunchKernelGGL(k_shared_scan, dim3(grid), dim3(kBlock), 0, s, d_q, d_items, nitems, d_counter, cp,
                     tile_rows);
  hip_check(hipGetLastError(), "k_shared_scan");
}

void launch_emit_count(const Emit* d_e, int nq, unsigned long long total_blocks, unsigned long long* d_counts,
                       unsigned long long* d_offsets, unsigned long long* d_totals, hipStream_t s) {
  if (nq <= 0) return;
  unsigned g = (unsigned)(total_blocks < 4096 ? (total_blocks ? total_blocks : 1) : 4096);
  hipLaunchKernelGGL(k_emit_count, dim3(g), dim3(kBlock), 0, s, d_e, nq, total_blocks, d_counts);
  hipLaunchKernelGGL(k_emit_scan, dim3(nq), dim3(kBlock), 0, s, d_e, nq, d_counts, d_offsets, d_totals);
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
