// Level-2 partition kernels (super-bucket regions -> buckets) and NN-Descent's
// reverse lists built with them. The layout and the staging are described in
// psq_partition.hpp.
#include "psq_kernels.hpp"
#include "psq_partition.hpp"

namespace psq {

// Level-2 count: block (s, c) takes the c-th of gridDim.y strided shares of
// super-bucket s's region [sstart[s], sstart[s+1]); its buckets are local ids
// 0 .. 2^(sshift - blog). Adds the counts to btot.
__global__ __launch_bounds__(1024) void k_part_l2_count(const uint32_t* __restrict__ sstart,
                                                        const uint32_t* __restrict__ pj1,
                                                        uint32_t sshift, uint32_t blog,
                                                        uint32_t* __restrict__ btot) {
  __shared__ uint32_t h[PART_BPS_MAX];
  const uint32_t tid = threadIdx.x, sb = blockIdx.x;
  const uint32_t nbk = 1u << (sshift - blog), bk0 = sb * nbk;
  for (uint32_t b = tid; b < nbk; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const uint32_t q0 = sstart[sb], q1 = sstart[sb + 1];
  const uint32_t stride = gridDim.y * PART_P2_CHUNK;
  for (uint32_t c0 = q0 + blockIdx.y * PART_P2_CHUNK; c0 < q1; c0 += stride)
    for (uint32_t q = c0 + tid; q < min(q1, c0 + PART_P2_CHUNK); q += blockDim.x)
      atomicAdd(&h[(pj1[q] >> blog) - bk0], 1u);
  __syncthreads();
  for (uint32_t b = tid; b < nbk; b += blockDim.x)
    if (h[b]) atomicAdd(&btot[bk0 + b], h[b]);
}

// Level-2 placement, same shares, chunk by chunk: pay2 = the entry key, pl2 =
// the target's row within its bucket, in bucket order.
__global__ __launch_bounds__(1024) void k_part_l2_scatter(
    const uint32_t* __restrict__ sstart, const uint64_t* __restrict__ pay1,
    const uint32_t* __restrict__ pj1, uint32_t sshift, uint32_t blog,
    const uint32_t* __restrict__ bstart, uint32_t* __restrict__ bfill, uint64_t* __restrict__ pay2,
    uint16_t* __restrict__ pl2) {
  extern __shared__ uint64_t s_key[];  // PART_P2_CHUNK keys, then PART_P2_CHUNK targets
  uint32_t* s_tgt = reinterpret_cast<uint32_t*>(s_key + PART_P2_CHUNK);
  __shared__ uint32_t s_cnt[PART_BPS_MAX + 1], s_base[PART_BPS_MAX], s_part[1024 / WAVE];
  const uint32_t sb = blockIdx.x;
  const uint32_t nbk = 1u << (sshift - blog), bk0 = sb * nbk;
  const uint32_t bmask = (1u << blog) - 1u;
  const uint32_t q0 = sstart[sb], q1 = sstart[sb + 1];
  const uint32_t stride = gridDim.y * PART_P2_CHUNK;
  for (uint32_t c0 = q0 + blockIdx.y * PART_P2_CHUNK; c0 < q1; c0 += stride) {
    const uint64_t* kin = pay1 + c0;
    const uint32_t* tin = pj1 + c0;
    stage_partition<PART_P2_CHUNK / 1024>(
        [&](uint32_t e, uint64_t& kk, uint32_t& tt) {
          kk = kin[e];
          tt = tin[e];
          return true;
        },
        min((uint32_t)PART_P2_CHUNK, q1 - c0), nbk, bk0,
        [&](uint32_t j) { return (j >> blog) - bk0; }, bstart, bfill, pay2, pl2,
        [&](uint32_t j) { return (uint16_t)(j & bmask); }, s_key, s_tgt, s_cnt, s_base, s_part);
  }
}

// ── reverse lists of NN-Descent (nndescent.rs:862-881) ──
// Every row entry G[e] (row i = e / k, neighbour j) becomes the reverse entry
// (class << 63) | (prio24 << 32) | i of row j, class 0 = new, 1 = old, with the
// priority hash of its slot e. The consumer (k_sample_select) is a set function
// of each row's reverse entries, so their order inside a row is free.

__device__ inline bool rv_load(const uint64_t* __restrict__ G, uint64_t e, uint32_t k,
                               uint64_t seed, uint64_t iter, uint64_t& key, uint32_t& j) {
  const uint64_t g = G[e];
  if (g == EMPTY_KEY) return false;
  j = key_idx(g);
  const uint64_t cls = (g & 1) ? 0ull : 1ull;
  key = (cls << 63) | ((uint64_t)hash_prio24(seed, iter, e) << 32) | (uint64_t)(e / k);
  return true;
}

// Level-1 count: non-empty entries per super-bucket of their neighbour.
__global__ __launch_bounds__(1024) void k_rv_l1_count(const uint64_t* __restrict__ G, uint64_t nk,
                                                      uint32_t sshift, uint32_t ns,
                                                      uint32_t* __restrict__ stot) {
  __shared__ uint32_t h[PART_SB_MAX];
  const uint32_t tid = threadIdx.x;
  if (tid < ns) h[tid] = 0;
  __syncthreads();
  const uint64_t e0 = (uint64_t)blockIdx.x * PART_C1_CHUNK;
  const uint64_t e1 = min(nk, e0 + PART_C1_CHUNK);
  for (uint64_t e = e0 + tid; e < e1; e += blockDim.x) {
    const uint64_t g = G[e];
    if (g != EMPTY_KEY) atomicAdd(&h[key_idx(g) >> sshift], 1u);
  }
  __syncthreads();
  if (tid < ns && h[tid]) atomicAdd(&stot[tid], h[tid]);
}

__global__ __launch_bounds__(1024) void k_rv_l1_scatter(
    const uint64_t* __restrict__ G, uint64_t nk, uint32_t k, uint64_t seed, uint64_t iter,
    uint32_t sshift, uint32_t ns, const uint32_t* __restrict__ sstart, uint32_t* __restrict__ sfill,
    uint64_t* __restrict__ pay1, uint32_t* __restrict__ pj1) {
  __shared__ uint64_t s_key[PART_P1_CHUNK];
  __shared__ uint32_t s_tgt[PART_P1_CHUNK];
  __shared__ uint32_t s_cnt[PART_SB_MAX + 1], s_base[PART_SB_MAX], s_part[1024 / WAVE];
  const uint64_t e0 = (uint64_t)blockIdx.x * PART_P1_CHUNK;
  const uint32_t cnt = (uint32_t)min((uint64_t)PART_P1_CHUNK, nk - e0);
  stage_partition<PART_P1_CHUNK / 1024>(
      [&](uint32_t e, uint64_t& kk, uint32_t& tt) {
        return rv_load(G, e0 + e, k, seed, iter, kk, tt);
      },
      cnt, ns, 0u, [&](uint32_t j) { return j >> sshift; }, sstart, sfill, pay1, pj1,
      [](uint32_t j) { return j; }, s_key, s_tgt, s_cnt, s_base, s_part);
}

// One block per bucket of 2^blog rows: counting sort of the bucket's entries by
// row into rev (the bucket's region), and the row offsets roff (roff[n] = the
// total). Dynamic LDS: 2 * 2^blog + 1 counters.
__global__ __launch_bounds__(1024) void k_rv_bucket_sort(const uint32_t* __restrict__ bstart,
                                                         const uint64_t* __restrict__ pay2,
                                                         const uint16_t* __restrict__ pl2, uint32_t n,
                                                         uint32_t blog, uint32_t nb,
                                                         uint32_t* __restrict__ roff,
                                                         uint64_t* __restrict__ rev) {
  extern __shared__ uint32_t s_c[];
  __shared__ uint32_t s_part[1024 / WAVE];
  const uint32_t tid = threadIdx.x, b = blockIdx.x;
  const uint32_t rows_b = 1u << blog, row0 = b << blog;
  const uint32_t rows = min(rows_b, n - row0);
  uint32_t* off = s_c;
  uint32_t* cur = s_c + rows_b + 1;
  const uint32_t p0 = bstart[b], p1 = bstart[b + 1], c = p1 - p0;
  for (uint32_t r = tid; r < rows; r += blockDim.x) off[r] = 0;
  __syncthreads();
  for (uint32_t t = tid; t < c; t += blockDim.x) atomicAdd(&off[pl2[p0 + t]], 1u);
  __syncthreads();
  block_scan4(off, rows, s_part);
  for (uint32_t r = tid; r < rows; r += blockDim.x) {
    cur[r] = off[r];
    roff[row0 + r] = p0 + off[r];
  }
  if (b == nb - 1 && tid == 0) roff[n] = p1;
  __syncthreads();
  for (uint32_t t = tid; t < c; t += blockDim.x)
    rev[p0 + atomicAdd(&cur[pl2[p0 + t]], 1u)] = pay2[p0 + t];
}

}  // namespace psq
