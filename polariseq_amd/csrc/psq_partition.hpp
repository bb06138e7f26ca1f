// LDS-staged partitioning of (key, target row) entries by target row range,
// shared by NN-Descent's reverse lists and the fuzzy connectivities' transpose.
//
// Entries go from their source order to groups of consecutive target rows in
// two passes, each counting in LDS and reserving its output runs with one
// global atomic per (block, part) instead of one per entry:
//   level 1: source chunks -> super-buckets of 2^sshift rows (at most
//            PART_SB_MAX of them): key array + full target row;
//   level 2: each super-bucket's region -> buckets of 2^blog rows: key array +
//            the target's row within its bucket (u16).
// Within a block, entries are sorted by part in LDS before they are written,
// so consecutive lanes write consecutive addresses of one part's run. The
// order of entries inside a part depends on LDS atomics: consumers must treat
// a part's entries as a set (or sort them).
#pragma once
#include "psq_device.hpp"

namespace psq {

constexpr int PART_SB_MAX = 64;      // most super-buckets
constexpr int PART_SB_WIDE = 1024;   // most super-buckets of the wide level-1 kernels
constexpr int PART_BPS_MAX = 4096;   // most buckets per super-bucket
constexpr int PART_C1_CHUNK = 65536; // entries per block of a level-1 count
constexpr int PART_P1_CHUNK = 4096;  // entries per block of a level-1 placement
constexpr int PART_P2_CHUNK = 8192;  // entries per chunk of the level-2 placement
constexpr int PART_L2_SHARES = 64;   // blocks per super-bucket in level 2

// Exclusive scan in place of v[0, L) by the whole block, four consecutive
// values per thread per round; v[L] = total. v may be in LDS or global memory.
__device__ inline void block_scan4(uint32_t* v, uint32_t L, uint32_t* s_part) {
  const uint32_t tid = threadIdx.x, l = tid & (WAVE - 1), wv = tid / WAVE;
  const uint32_t nwv = blockDim.x / WAVE;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < L; base += 4 * blockDim.x) {
    const uint32_t i0 = base + 4 * tid;
    uint32_t x[4], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      x[k] = i0 + k < L ? v[i0 + k] : 0u;
      sum += x[k];
    }
    uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, WAVE);
      if (l >= (uint32_t)o) inc += y;
    }
    if (l == WAVE - 1) s_part[wv] = inc;
    __syncthreads();
    if (wv == 0) {
      uint32_t t = l < nwv ? s_part[l] : 0u;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(t, o, WAVE);
        if (l >= (uint32_t)o) t += y;
      }
      if (l < nwv) s_part[l] = t;
    }
    __syncthreads();
    uint32_t run = carry + (wv > 0 ? s_part[wv - 1] : 0u) + inc - sum;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (i0 + k < L) {
        v[i0 + k] = run;
        run += x[k];
      }
    carry += s_part[nwv - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) v[L] = carry;
  __syncthreads();
}

// One staged partition of the entries e < cnt that load(e, key, tgt) accepts
// (at most NPT * blockDim.x) into parts part(tgt) < np: counted in LDS, each
// part's run reserved in global memory with one atomic (base = start[part0 + p]
// + atomicAdd(fill[part0 + p], count)), sorted by part in LDS, then written so
// that consecutive threads write consecutive addresses of one run. The second
// output array receives out_jv(tgt). Every thread of the block calls it.
template <int NPT, typename Load, typename PartF, typename OutJ, typename JT>
__device__ inline void stage_partition(Load load, uint32_t cnt, uint32_t np, uint32_t part0,
                                       PartF part, const uint32_t* __restrict__ start,
                                       uint32_t* __restrict__ fill, uint64_t* __restrict__ out_key,
                                       JT* __restrict__ out_j, OutJ out_jv, uint64_t* s_key,
                                       uint32_t* s_tgt, uint32_t* s_cnt, uint32_t* s_base,
                                       uint32_t* s_part) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t p = tid; p < np; p += blockDim.x) s_cnt[p] = 0;
  __syncthreads();
  uint64_t k[NPT];
  uint32_t t[NPT], rank[NPT];
  bool ok[NPT];
#pragma unroll
  for (int q = 0; q < NPT; ++q) {
    const uint32_t e = q * blockDim.x + tid;
    ok[q] = e < cnt && load(e, k[q], t[q]);
    if (ok[q]) rank[q] = atomicAdd(&s_cnt[part(t[q])], 1u);
  }
  __syncthreads();
  for (uint32_t p = tid; p < np; p += blockDim.x) {
    const uint32_t c = s_cnt[p];
    s_base[p] = c ? start[part0 + p] + atomicAdd(&fill[part0 + p], c) : 0u;
  }
  block_scan4(s_cnt, np, s_part);  // s_cnt becomes the LDS offsets, s_cnt[np] the total
  const uint32_t total = s_cnt[np];
#pragma unroll
  for (int q = 0; q < NPT; ++q)
    if (ok[q]) {
      const uint32_t pos = s_cnt[part(t[q])] + rank[q];
      s_key[pos] = k[q];
      s_tgt[pos] = t[q];
    }
  __syncthreads();
  for (uint32_t e = tid; e < total; e += blockDim.x) {
    const uint32_t tg = s_tgt[e], p = part(tg);
    const uint32_t pos = s_base[p] + (e - s_cnt[p]);
    out_key[pos] = s_key[e];
    out_j[pos] = out_jv(tg);
  }
  __syncthreads();
}

// Super-bucket shift and bucket size for n target rows: super-buckets of
// 2^sshift rows, at most sb_max (PART_SB_MAX, or PART_SB_WIDE for the wide
// level-1 kernels); buckets of 2^blog rows, blog >= blog_min and at most
// PART_BPS_MAX buckets per super-bucket.
struct PartShape {
  uint32_t sshift, ns, blog, nb;
};
__host__ inline PartShape part_shape(uint64_t n, uint32_t blog_min, uint32_t sb_max = PART_SB_MAX) {
  PartShape p;
  p.sshift = blog_min;
  while (((n + (1ull << p.sshift) - 1) >> p.sshift) > (uint64_t)sb_max) ++p.sshift;
  p.blog = p.sshift > blog_min + 12 ? p.sshift - 12 : blog_min;
  p.ns = (uint32_t)((n + (1ull << p.sshift) - 1) >> p.sshift);
  p.nb = (uint32_t)((n + (1ull << p.blog) - 1) >> p.blog);
  return p;
}

}  // namespace psq
