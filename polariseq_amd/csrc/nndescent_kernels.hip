// NN-Descent kernels for gfx950 (wave64): random fill, reverse-neighbour
// sampling, local join, per-row merge and CSR extraction.
//
// The algorithm is crates/polariseq-core/src/nndescent.rs:699-1003, rebuilt so
// every per-row step is a set function of its inputs and the graph does not
// depend on thread order:
//   * rows are sorted u64 keys (dist_bits << 32 | idx << 1 | new);
//   * sampling selects, per row and flag class, the m smallest (priority, nbr)
//     keys among forward and reverse proposals (nndescent.rs:853-909);
//   * the join (nndescent.rs:912-969) routes each proposal to one 8-slot bucket
//     of its destination row. An atomic-min cascade keeps the bucket's 8
//     smallest distinct keys whatever the arrival order;
//   * the merge takes top-k of (row ∪ buckets), deduplicated on (dist, idx).
// oracle/gpu_model.cpp states the same semantics on the CPU; the GPU tests
// compare the two graphs bit for bit.
#include "psq_kernels.hpp"

namespace psq {

// ── small utilities ──
__global__ void k_fill_u64(uint64_t* p, size_t count, uint64_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, size_t count, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}
// Sums the counter stripes: out[w] = sum over stripes of ctr[stripe][w].
__global__ void k_ctr_reduce(const unsigned long long* __restrict__ ctr,
                             unsigned long long* __restrict__ out) {
  const uint32_t w = threadIdx.x;
  if (w >= CTR_WORDS) return;
  unsigned long long t = 0;
  for (uint32_t st = 0; st < CTR_STRIPES; ++st) t += ctr[(size_t)st * CTR_WORDS + w];
  out[w] = t;
}
// Adds the join's per-cell records (words 0, 1, 3, 4) to the counter stripes: a
// grid-stride pass of coalesced 16-byte reads, then one atomic per wave and word.
__global__ __launch_bounds__(256) void k_join_rec_reduce(const uint32_t* __restrict__ rec, uint32_t n,
                                                         unsigned long long* __restrict__ counters) {
  uint64_t a0 = 0, a1 = 0, a3 = 0, a4 = 0;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    const uint4 lo = *reinterpret_cast<const uint4*>(rec + i * JOIN_REC_WORDS);
    const uint32_t hi = rec[i * JOIN_REC_WORDS + 4];
    a0 += lo.x;
    a1 += lo.y;
    a3 += lo.w;
    a4 += hi;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a0 += shfl_xor64(a0, o);
    a1 += shfl_xor64(a1, o);
    a3 += shfl_xor64(a3, o);
    a4 += shfl_xor64(a4, o);
  }
  if (lane_id() == 0) {
    ctr_add(counters, 0, a0);
    ctr_add(counters, 1, a1);
    ctr_add(counters, 3, a3);
    ctr_add(counters, 4, a4);
  }
}

// counters[stripe][word] += the sum of v[0, n): a grid-stride pass, then one
// atomic per wave.
__global__ __launch_bounds__(256) void k_sum_u32_to(const uint32_t* __restrict__ v, uint32_t n,
                                                    unsigned long long* __restrict__ counters,
                                                    uint32_t word) {
  uint64_t a = 0;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) a += v[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += shfl_xor64(a, o);
  if (lane_id() == 0 && a) ctr_add(counters, word, a);
}

__global__ void k_fill_u8(uint8_t* p, size_t count, uint8_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// Copy rows of d floats into rows of dp floats (zero padded), so each row starts
// on a 128-byte boundary and can be read with 16-byte loads.
__global__ void k_pad_rows(const float* __restrict__ X, uint32_t n, uint32_t d, uint32_t dp,
                           float* __restrict__ Xp) {
  size_t total = (size_t)n * dp;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    size_t r = i / dp, c = i - r * dp;
    Xp[i] = c < d ? X[r * d + c] : 0.0f;
  }
}

// ── exclusive scan of u32 counts (three passes) ──
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  // blockDim.x == SCAN_BLOCK threads; wave-level inclusive scan then across waves.
  const int l = lane_id();
  const int w = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(x, o, WAVE);
    if (l >= o) x += y;
  }
  if (l == WAVE - 1) sh[w] = x;
  __syncthreads();
  if (w == 0) {
    const int nw = SCAN_BLOCK / WAVE;
    uint32_t s = l < nw ? sh[l] : 0u;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      uint32_t y = __shfl_up(s, o, WAVE);
      if (l >= o) s += y;
    }
    if (l < nw) sh[l] = s;
  }
  __syncthreads();
  uint32_t before = (w > 0 ? sh[w - 1] : 0u) + (x - v);
  total = sh[SCAN_BLOCK / WAVE - 1];
  __syncthreads();
  return before;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_local(const uint32_t* __restrict__ in,
                                                           uint32_t* __restrict__ out, uint32_t n,
                                                           uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t sh[SCAN_BLOCK / WAVE];
  uint32_t i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
  uint32_t v = i < n ? in[i] : 0u;
  uint32_t total;
  uint32_t ex = block_exclusive_scan(v, sh, total);
  if (i < n) out[i] = ex;
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Scans the block sums in place (one block, loops over chunks) and writes the
// grand total to *total_out.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nb,
                                                          uint32_t* __restrict__ total_out) {
  __shared__ uint32_t sh[SCAN_BLOCK / WAVE];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += SCAN_BLOCK) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t total;
    uint32_t ex = block_exclusive_scan(v, sh, total);
    if (i < nb) sums[i] = ex + carry;
    carry += total;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, uint32_t n,
                           const uint32_t* __restrict__ sums) {
  uint32_t i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
  if (i < n) out[i] += sums[blockIdx.x];
}

// ── random fill (nndescent.rs:819-845) ──
// One thread per row. A row that still holds EMPTY takes hash_index draws in
// attempt order, skipping its own index (j == i -> j + 1) and indices already
// present, until it is full or 64 draws are spent. New keys are appended
// unsorted; the row is marked so the next merge pass sorts it.
__global__ void k_random_fill(const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp,
                              uint64_t* __restrict__ G, uint32_t k, uint64_t seed,
                              uint8_t* __restrict__ touched) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t* r = G + (size_t)i * k;
  if (r[k - 1] != EMPTY_KEY) return;
  uint32_t cnt = 0;
  while (cnt < k && r[cnt] != EMPTY_KEY) ++cnt;
  const float* xi = Xp + (size_t)i * dp;
  bool wrote = false;
  for (uint64_t attempt = 0; cnt < k && attempt < 64; ++attempt) {
    uint64_t j = hash_index(seed, i, attempt, n);
    if (j == i) j = (j + 1) % n;
    bool dup = false;
    for (uint32_t s = 0; s < cnt; ++s)
      if (key_idx(r[s]) == (uint32_t)j) {
        dup = true;
        break;
      }
    if (dup) continue;
    float dd = sq_dist_ref(xi, Xp + j * dp, (int)d);
    if (dd >= __builtin_inff()) continue;
    r[cnt++] = make_key(dd, (uint32_t)j, 1);
    wrote = true;
  }
  if (wrote) touched[i] = 1;
}

// One point (nndescent.rs:727, 836-839): k is 1, the random fill's only draw
// is the point itself (hash_index % 1 = 0, (0 + 1) % 1 = 0), pushed with its own
// distance into the empty row, and no later step changes the row.
__global__ void k_single_point(const float* __restrict__ X, uint32_t d, uint32_t* __restrict__ indptr,
                               uint32_t* __restrict__ indices, float* __restrict__ dists) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  indptr[0] = 0;
  indptr[1] = 1;
  indices[0] = 0;
  dists[0] = (float)__dsqrt_rn((double)sq_dist_ref(X, X, (int)d));
}

// ── sampling (nndescent.rs:853-909) ──
// One wave (one block) per row i. Candidates: the row's own entries (forward,
// priority of slot i*k+s) and the reverse entries rev[roff[i]..roff[i+1]).
// Per flag class, a neighbour j seen twice keeps the copy from the earlier
// source row (the forward one when i < j; the reference applies candidate
// updates in source order and rejects later duplicates, nndescent.rs:244-249,
// 856-892); then the m smallest (priority, neighbour) keys form the new
// and old lists. Entries of the row that made the new list lose their new flag
// (nndescent.rs:894-909). Candidate t < k is forward entry t, t >= k reverse
// entry t - k; they are taken 64 at a time. Key: (class << 63) | (prio24 << 32)
// | neighbour, class 0 = new, so one sort orders new before old.
// The row's neighbours are also pinned into its buckets for this iteration's
// join (at most BUCKET per bucket, in any order; the other slots empty), and
// the row's threshold for the join is set.
// KV: forward vectors, k <= 64 KV and S <= 256 KV; lane l holds forward entries
// l, l + 64, ..., l + 64 (KV - 1).
template <int KV>
__global__ __launch_bounds__(64) void k_sample_select(
    uint64_t* __restrict__ G, uint32_t n, uint32_t k, uint32_t m, uint64_t seed, uint64_t iter,
    const uint32_t* __restrict__ roff, const uint64_t* __restrict__ rev,
    uint32_t* __restrict__ newl, uint32_t* __restrict__ oldl, uint32_t* __restrict__ ncnt,
    uint64_t* __restrict__ slots, uint32_t S, float* __restrict__ thr, uint32_t thr_rank) {
  constexpr int KF = KV * WAVE;  // forward entries
  constexpr int HB = KV == 1 ? 7 : (KV == 2 ? 8 : 9);  // hash table of 2 KF slots
  constexpr uint32_t HM = (1u << HB) - 1;
  __shared__ uint64_t s_fk[KF];            // forward keys
  __shared__ uint64_t s_kkey[KF];          // key of the reverse entry that replaced a forward one
  __shared__ uint64_t s_pin[256 * KV];     // bucket layout
  __shared__ uint32_t s_bcnt[32 * KV];     // pinned per bucket
  __shared__ uint32_t s_ht[2 * KF];        // forward neighbour -> entry + 1 (open addressing)
  const uint32_t l = (uint32_t)lane_id();
  const uint32_t i = block_id();
  if (i >= n) return;  // block-uniform
  // the reverse list's bounds, loaded together with the row
  const uint32_t r0 = roff[i], r1 = roff[i + 1];
  uint64_t fkey[KV], fk[KV];
  bool fvalid[KV];
  uint32_t me[KV];
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    const uint32_t f = v * WAVE + l;
    fkey[v] = f < k ? G[(size_t)i * k + f] : EMPTY_KEY;
    fvalid[v] = fkey[v] != EMPTY_KEY;
    me[v] = key_idx(fkey[v]);
    // the join's threshold: distance of entry thr_rank - 1 (the row's worst when
    // thr_rank = k), infinite while the row holds fewer entries
    if (f == thr_rank - 1) thr[i] = fvalid[v] ? key_dist(fkey[v]) : __builtin_inff();
    fk[v] = EMPTY_KEY;
    if (fvalid[v]) {
      const uint64_t cls = (fkey[v] & 1) ? 0ull : 1ull;
      const uint64_t pr = hash_prio24(seed, iter, (uint64_t)i * k + f);
      fk[v] = (cls << 63) | (pr << 32) | me[v];
    }
    s_fk[f] = fk[v];
    s_kkey[f] = EMPTY_KEY;
  }
  for (uint32_t e = l; e < 32 * KV; e += WAVE) s_bcnt[e] = 0;
  for (uint32_t e = l; e < 2 * KF; e += WAVE) s_ht[e] = 0;
  __syncthreads();
#pragma unroll
  for (int v = 0; v < KV; ++v)
    if (fvalid[v]) {  // forward neighbours are distinct
      for (uint32_t h = (me[v] * 0x9E3779B1u) >> (32 - HB);; h = (h + 1) & HM)
        if (atomicCAS(&s_ht[h], 0u, v * WAVE + l + 1) == 0u) break;
    }
  __syncthreads();
  const uint32_t nrev = r1 - r0, T = k + nrev;
  const uint32_t nvec = (T + WAVE - 1) / WAVE;
  // Reverse entry of candidate t (t >= k; vectors 0 and 1 are kept in registers
  // after their first read).
  uint64_t rvc0 = 0, rvc1 = 0;
  auto rev_entry = [&](uint32_t j, uint32_t t) -> uint64_t {
    if (j == 0) return rvc0;
    if (j == 1) return rvc1;
    return rev[r0 + (t - k)];
  };
  // Reverse candidate t = j WAVE + l (t >= k): its key, or EMPTY when a forward
  // entry of the same class names the same neighbour and comes from the earlier
  // source row (i before the neighbour). When the reverse copy's source (the
  // neighbour) comes first, it replaces that forward entry (record = true).
  auto rev_key = [&](uint32_t j, bool record) -> uint64_t {
    const uint32_t t = j * WAVE + l;
    if (t < k || t >= T) return EMPTY_KEY;
    const uint64_t rv = rev_entry(j, t);
    const uint32_t src = (uint32_t)rv;
    const uint64_t cp = rv >> 32;  // class and priority
    // the forward entry with this neighbour, if any (forward neighbours are distinct)
    for (uint32_t h = (src * 0x9E3779B1u) >> (32 - HB);; h = (h + 1) & HM) {
      const uint32_t slot = s_ht[h];
      if (slot == 0) return rv;
      const uint32_t f = slot - 1;
      const uint64_t fv = s_fk[f];
      if ((uint32_t)fv != src) continue;
      if ((fv >> 63) != (cp >> 31)) return rv;  // other class: both stay
      if (src < i) {
        if (record) s_kkey[f] = rv;
        return rv;
      }
      return EMPTY_KEY;
    }
  };
  // pass 1: every replacement is known before a forward key is used
  if (nvec > 0 && l >= k && l < T) rvc0 = rev[r0 + (l - k)];
  if (nvec > 1 && WAVE + l >= k && WAVE + l < T) rvc1 = rev[r0 + (WAVE + l - k)];
  for (uint32_t j = 0; j < nvec; ++j) (void)rev_key(j, true);
  __syncthreads();
  // pass 2: selection. A row's lists are sets: the join pairs their members in
  // any order (every proposal goes through the order-free bucket insert), so a
  // class with at most m entries needs no sort; otherwise its m smallest keys
  // (class, priority, neighbour) are kept.
  auto key_of = [&](uint32_t j) -> uint64_t {
    const uint32_t t = j * WAVE + l;
    if (t < k) {
      uint64_t f = EMPTY_KEY;
      bool fv = false;
#pragma unroll
      for (int v = 0; v < KV; ++v)
        if ((uint32_t)v == j) {
          f = fk[v];
          fv = fvalid[v];
        }
      return (fv && s_kkey[t] == EMPTY_KEY) ? f : EMPTY_KEY;
    }
    return rev_key(j, false);
  };
  uint32_t nn = 0, no = 0;
  uint64_t lastN = 0;  // the largest new-class key kept
  bool sorted = true;
  if (nvec == 1) {
    const uint64_t key = key_of(0);
    const bool isN = key != EMPTY_KEY && (key >> 63) == 0;
    const bool isO = key != EMPTY_KEY && (key >> 63) != 0;
    const uint64_t mN = __ballot(isN), mO = __ballot(isO);
    const uint32_t cN = (uint32_t)__popcll(mN), cO = (uint32_t)__popcll(mO);
    if (cN <= m && cO <= m) {  // uniform: every entry is kept
      const uint64_t below = (1ull << l) - 1ull;
      if (isN) newl[(size_t)i * m + (uint32_t)__popcll(mN & below)] = (uint32_t)key;
      if (isO) oldl[(size_t)i * m + (uint32_t)__popcll(mO & below)] = (uint32_t)key;
      if (l >= cN && l < m) newl[(size_t)i * m + l] = EMPTY;
      if (l >= cO && l < m) oldl[(size_t)i * m + l] = EMPTY;
      nn = cN;
      no = cO;
      lastN = ~0ull >> 1;  // above every new-class key
      sorted = false;
    }
  }
  if (sorted) {
    uint64_t bestN = EMPTY_KEY, bestO = EMPTY_KEY;
    for (uint32_t j = 0; j < nvec; ++j) {
      const uint64_t sk = wave_sort64(key_of(j));
      const uint32_t cN = (uint32_t)__popcll(__ballot(sk != EMPTY_KEY && (sk >> 63) == 0));
      const uint32_t cO = (uint32_t)__popcll(__ballot(sk != EMPTY_KEY && (sk >> 63) == 1));
      const uint64_t vN = l < cN ? sk : EMPTY_KEY;
      const uint64_t sh = shfl64(sk, (int)((cN + l) & (WAVE - 1)));
      const uint64_t vO = l < cO ? sh : EMPTY_KEY;
      if (j == 0) {
        bestN = vN;
        bestO = vO;
      } else {
        bestN = wave_merge_keep64(bestN, vN);
        bestO = wave_merge_keep64(bestO, vO);
      }
      nn += cN;
      no += cO;
    }
    nn = nn < m ? nn : m;
    no = no < m ? no : m;
    if (l < m) {
      newl[(size_t)i * m + l] = l < nn ? (uint32_t)bestN : EMPTY;
      oldl[(size_t)i * m + l] = l < no ? (uint32_t)bestO : EMPTY;
    }
    lastN = shfl64(bestN, (int)(nn > 0 ? nn - 1 : 0));
  }
  if (l == 0) ncnt[i] = nn | (no << 16);
  // flag update: the surviving new-class key of this neighbour (the entry's own
  // or the reverse entry that replaced it) is among the nn selected
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    const uint32_t f = v * WAVE + l;
    if (fvalid[v] && (fkey[v] & 1) && nn > 0) {
      const uint64_t kk = s_kkey[f] != EMPTY_KEY ? s_kkey[f] : fk[v];
      if (kk <= lastN) G[(size_t)i * k + f] = fkey[v] & ~1ull;
    }
  }
  // bucket layout
  const SlotHash sh = slot_hash_of(seed, iter, S);
  for (uint32_t e = l; e < S; e += WAVE) s_pin[e] = EMPTY_KEY;
  __syncthreads();
#pragma unroll
  for (int v = 0; v < KV; ++v)
    if (fvalid[v]) {
      const uint32_t b = slot_bucket(sh, me[v]);
      const uint32_t r = atomicAdd(&s_bcnt[b], 1u);
      if (r < BUCKET) s_pin[b * BUCKET + r] = pin_key(me[v]);
    }
  __syncthreads();
  for (uint32_t e = l; e < S; e += WAVE) slots[(size_t)i * S + e] = s_pin[e];
}
template __global__ void k_sample_select<1>(uint64_t*, uint32_t, uint32_t, uint32_t, uint64_t, uint64_t,
                                            const uint32_t*, const uint64_t*, uint32_t*, uint32_t*,
                                            uint32_t*, uint64_t*, uint32_t, float*, uint32_t);
template __global__ void k_sample_select<2>(uint64_t*, uint32_t, uint32_t, uint32_t, uint64_t, uint64_t,
                                            const uint32_t*, const uint64_t*, uint32_t*, uint32_t*,
                                            uint32_t*, uint64_t*, uint32_t, float*, uint32_t);
template __global__ void k_sample_select<4>(uint64_t*, uint32_t, uint32_t, uint32_t, uint64_t, uint64_t,
                                            const uint32_t*, const uint64_t*, uint32_t*, uint32_t*,
                                            uint32_t*, uint64_t*, uint32_t, float*, uint32_t);

// ── local join (nndescent.rs:912-969) ──
// One wave (one block) per cell i. The candidate rows (new list, then old
// list) are staged in LDS with coalesced 16-byte loads, at a row stride RS
// (floats, RS / 4 odd) that spreads rows over the LDS banks. The cell's pairs,
// new x new (a < b) then new x old with different points, are numbered and
// dealt to the 64 lanes, so every lane computes a distance per round. A pair
// whose distance beats an endpoint's threshold (snapshot of that row's worst)
// proposes the other endpoint to it; proposals are batched through the wave's
// LDS list. D = 50 fixes the dimension at compile time; D = 0 reads it at run
// time. Both use the reference's operation order (simd.rs:45-62). A block holds
// R candidate rows; a cell with more is left to a launch with more (jc).
template <int D>
__global__ __launch_bounds__(64) void k_join(
    const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp, uint32_t RS,
    const uint32_t* __restrict__ newl, const uint32_t* __restrict__ oldl,
    const uint32_t* __restrict__ ncnt, uint32_t m, uint32_t R, JoinCells jc,
    const float* __restrict__ thr, uint64_t* __restrict__ slots, uint8_t* __restrict__ touched,
    uint32_t S, uint64_t seed, uint64_t iter, uint32_t* __restrict__ rec) {
  // LDS: pb_key[PB_CAP] u64 | rows[R][RS] f32 | ids[R] | thr[R] | pb_row[PB_CAP]
  extern __shared__ __attribute__((aligned(16))) float lds[];
  uint64_t* pb_key = reinterpret_cast<uint64_t*>(lds);
  float* rows = lds + 2 * PB_CAP;
  uint32_t* s_id = reinterpret_cast<uint32_t*>(rows + R * RS);
  float* s_thr = reinterpret_cast<float*>(s_id + R);
  uint32_t* pb_row = reinterpret_cast<uint32_t*>(s_thr + R);
  const uint32_t l = (uint32_t)lane_id();
  const SlotHash sh = slot_hash_of(seed, iter, S);
  uint64_t pairs = 0, sent = 0, rows_in = 0;
  uint32_t i;
  if (!join_cell_of_block(jc, n, i)) return;
  [&] {
    uint32_t nn, no;
    if (!stage_cell<D>(Xp, d, dp, RS, newl, oldl, ncnt, m, thr, i, R, s_id, s_thr, rows,
                       nullptr, nn, no))
      return;
    rows_in += nn + no;
    __syncthreads();
    const uint32_t P1 = nn * (nn - 1) / 2, P = P1 + nn * no;
    uint32_t pcnt = 0;
    for (uint32_t base = 0; base < P; base += WAVE) {
      const uint32_t t = base + l;
      bool pa = false, pb = false;
      uint32_t ia = 0, ib = 0;
      float dist = 0.0f;
      if (t < P) {
        uint32_t a, b;
        pair_decode(t, nn, no, P1, a, b);
        ia = s_id[a];
        ib = s_id[b];
        if (ia != ib) {
          ++pairs;
          dist = lds_sq_dist<D>(rows + __umul24(a, RS), rows + __umul24(b, RS), d);
          pa = dist < s_thr[a];
          pb = dist < s_thr[b];
        }
      }
      pb_append(pb_row, pb_key, pcnt, pa, ia, make_key(dist, ib, 1), sent);
      pb_flush(pb_row, pb_key, pcnt, false, slots, touched, S, sh);
      pb_append(pb_row, pb_key, pcnt, pb, ib, make_key(dist, ia, 1), sent);
      pb_flush(pb_row, pb_key, pcnt, false, slots, touched, S, sh);
    }
    pb_flush(pb_row, pb_key, pcnt, true, slots, touched, S, sh);
  }();
  // the cell's record: pairs, candidate rows gathered, proposals sent (every
  // pair is exact here; word 3 stays zero)
  uint32_t tot = (uint32_t)pairs;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, WAVE);
  join_record(rec, i, l, tot, (uint32_t)rows_in, 0u, (uint32_t)sent);
}

// Cells of one deferred join launch: flag[i] = 1 when cell i has a new
// candidate and lo < candidates <= hi.
__global__ void k_join_defer_flags(const uint32_t* __restrict__ ncnt, uint32_t n, uint32_t lo,
                                   uint32_t hi, uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = ncnt[i], nn = c & 0xFFFF, nc = nn + (c >> 16);
  flag[i] = (nn > 0 && nc > lo && nc <= hi) ? 1u : 0u;
}
// The flagged cells in cell order: list[pos[i]] = i (pos = exclusive scan of flag).
__global__ void k_join_defer_list(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                  uint32_t n, uint32_t* __restrict__ list) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) list[pos[i]] = i;
}

// ── merge (the apply of nndescent.rs:307-324 as a set function) ──
// One wave (one block) per touched row (every row when `force`): the row
// becomes the top k of (row ∪ proposals). Word 2 of the striped counters
// `changed` counts entries of the new row absent from the old.
// Proposals never repeat a neighbour of the row (pinned slots reject those)
// and are distinct (a neighbour has one bucket, which never holds a key twice),
// so when they fit one 64-lane vector the merge is: compact them, sort,
// and keep the 64 smallest of (row, proposals) with one bitonic merge (the row
// is sorted, except before the forced first merge, which sorts it). Otherwise
// every slot and the row are sorted together and deduplicated on (dist, idx),
// an entry already in the row keeping its flag.
template <int NR, int NT>
__global__ __launch_bounds__(64) void k_merge(uint64_t* __restrict__ G, uint32_t n, uint32_t k,
                                              uint64_t* __restrict__ slots, uint32_t S,
                                              uint8_t* __restrict__ touched, int force,
                                              uint32_t* __restrict__ chg) {
  __shared__ uint64_t s_prop[WAVE];
  const uint32_t l = (uint32_t)lane_id();
  const uint32_t p = block_id();
  if (p >= n) return;
  if (!force && !touched[p]) {
    if (l == 0) chg[p] = 0u;
    return;
  }
  uint64_t* row = G + (size_t)p * k;
  const uint64_t* sl = slots + (size_t)p * S;
  const uint64_t orow = l < k ? row[l] : EMPTY_KEY;
  if (force && !touched[p]) {
    // the forced pass only sorts (no proposals exist yet): the leaf pairs
    // leave rows sorted and the random fill marks the rows it extends, so a
    // sorted row is left as it is
    const uint64_t next = shfl64(orow, (int)((l + 1) & (WAVE - 1)));
    if (!__ballot(l + 1 < k && orow > next)) {
      if (l == 0) chg[p] = 0u;
      return;
    }
  }
  if (l == 0) touched[p] = 0;
  uint64_t v[NR];
  uint32_t total = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    v[r] = bucket_to_row(sl[r * WAVE + l]);
    const bool has = v[r] != EMPTY_KEY;
    const uint64_t mask = __ballot(has);
    const uint32_t pos = total + (uint32_t)__popcll(mask & ((1ull << l) - 1ull));
    if (has && pos < (uint32_t)WAVE) s_prop[pos] = v[r];
    total += (uint32_t)__popcll(mask);
  }
  __syncthreads();
  uint32_t nchanged = 0;
  if (total <= (uint32_t)WAVE) {
    const uint64_t pv = l < total ? s_prop[l] : EMPTY_KEY;
    const uint64_t sr = force ? wave_sort64(orow) : orow;
    const uint64_t mg = wave_merge_keep64(sr, wave_sort_prefix(pv, total));
    const uint32_t nall = (uint32_t)__popcll(__ballot(mg != EMPTY_KEY));
    const uint32_t nk = nall < k ? nall : k;
    const uint64_t kth = shfl64(mg, (int)(nk > 0 ? nk - 1 : 0));
    const uint32_t nold =
        nk > 0 ? (uint32_t)__popcll(__ballot(orow != EMPTY_KEY && orow <= kth)) : 0u;
    if (l < k) row[l] = l < nk ? mg : EMPTY_KEY;
    if (l == 0) nchanged = nk - nold;
  } else {
    uint64_t u[NT];
#pragma unroll
    for (int r = 0; r < NT; ++r) u[r] = r < NR ? v[r] : (r == NR ? orow : EMPTY_KEY);
    wave_sort_n<NT>(u);
    // dedup on (dist, idx): drop an element equal to its predecessor's
    uint32_t before = 0;
    uint64_t lastkey = EMPTY_KEY;
#pragma unroll
    for (int r = 0; r < NT; ++r) {
      uint64_t prev = shfl64(u[r], (int)((l + WAVE - 1) & (WAVE - 1)));
      if (l == 0) prev = (r > 0) ? lastkey : EMPTY_KEY;
      const bool has_prev = (r > 0) || (l > 0);
      const bool keep = u[r] != EMPTY_KEY && !(has_prev && (u[r] >> 1) == (prev >> 1));
      const uint64_t mask = __ballot(keep);
      const uint32_t rank = before + (uint32_t)__popcll(mask & ((1ull << l) - 1ull));
      bool found = false;
      for (uint32_t s2 = 0; s2 < k; ++s2) {
        const uint64_t o = shfl64(orow, (int)s2);
        found |= (o != EMPTY_KEY && (o >> 1) == (u[r] >> 1));
      }
      if (keep && rank < k) {
        row[rank] = u[r];
        nchanged += !found;
      }
      before += (uint32_t)__popcll(mask);
      lastkey = shfl64(u[r], WAVE - 1);
    }
    const uint32_t nkept = before < k ? before : k;
    for (uint32_t pos = nkept + l; pos < k; pos += WAVE) row[pos] = EMPTY_KEY;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nchanged += __shfl_xor(nchanged, o, WAVE);
  }
  if (l == 0) chg[p] = nchanged;
}

// Merge for rows of more than 64 entries or more than 256 slots: one block of
// 256 threads per touched row sorts the row and every slot together in LDS
// (bitonic, P >= k + S keys), keeps the first of each run of equal (dist, idx)
// (an entry already in the row keeps its flag, its key being the smaller), and
// writes the first k; the same set function as k_merge.
template <int P>
__global__ __launch_bounds__(256) void k_merge_wide(uint64_t* __restrict__ G, uint32_t n, uint32_t k,
                                                    uint64_t* __restrict__ slots, uint32_t S,
                                                    uint8_t* __restrict__ touched, int force,
                                                    uint32_t* __restrict__ chg) {
  constexpr int NTH = 256, EPT = P / NTH;
  __shared__ uint64_t s_key[P];
  __shared__ uint32_t s_wsum[NTH / WAVE];
  __shared__ uint64_t s_kth;
  const uint32_t tid = threadIdx.x, l = (uint32_t)lane_id(), w = tid / WAVE;
  const uint32_t p = block_id();
  if (p >= n) return;  // block-uniform
  const bool was = touched[p] != 0;
  if (!force && !was) {
    if (tid == 0) chg[p] = 0u;
    return;
  }
  uint64_t* row = G + (size_t)p * k;
  const uint64_t* sl = slots + (size_t)p * S;
  const uint64_t orow = tid < k ? row[tid] : EMPTY_KEY;  // k <= 256
  if (force && !was) {
    // the forced pass only sorts: a sorted row is left as it is
    const bool desc = tid + 1 < k && orow > row[tid + 1];
    if (!__syncthreads_or(desc)) {
      if (tid == 0) chg[p] = 0u;
      return;
    }
  }
  __syncthreads();
  if (tid == 0) touched[p] = 0;
  for (uint32_t e = tid; e < P; e += NTH)
    s_key[e] = e < k ? row[e] : (e < k + S ? bucket_to_row(sl[e - k]) : EMPTY_KEY);
  for (uint32_t size = 2; size <= P; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (uint32_t e = tid; e < P / 2; e += NTH) {
        const uint32_t lo = 2 * stride * (e / stride) + (e % stride), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t a = s_key[lo], b = s_key[hi];
        if ((a > b) == up) {
          s_key[lo] = b;
          s_key[hi] = a;
        }
      }
    }
  __syncthreads();
  // dedup on (dist, idx) and compact: thread tid takes elements [tid EPT, (tid + 1) EPT)
  bool keep[EPT];
  uint32_t mine = 0;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const uint32_t e = tid * EPT + q;
    const uint64_t key = s_key[e];
    keep[q] = key != EMPTY_KEY && (e == 0 || (key >> 1) != (s_key[e - 1] >> 1));
    mine += keep[q];
  }
  uint32_t incl = mine;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, WAVE);
    if (l >= (uint32_t)o) incl += y;
  }
  if (l == WAVE - 1) s_wsum[w] = incl;
  __syncthreads();
  uint32_t base = 0, total = 0;
  for (uint32_t q = 0; q < NTH / WAVE; ++q) {
    base += q < w ? s_wsum[q] : 0u;
    total += s_wsum[q];
  }
  uint32_t rank = base + incl - mine;
  const uint32_t nk = total < k ? total : k;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    if (keep[q]) {
      const uint64_t key = s_key[tid * EPT + q];
      if (rank < k) row[rank] = key;
      if (rank + 1 == nk) s_kth = key;
      ++rank;
    }
  }
  for (uint32_t pos = nk + tid; pos < k; pos += NTH) row[pos] = EMPTY_KEY;
  __syncthreads();
  // entries of the new row absent from the old: nk minus the old entries kept
  const uint64_t kth = s_kth;
  const bool kept_old = nk > 0 && orow != EMPTY_KEY && (orow >> 1) <= (kth >> 1);
  const uint32_t nold = (uint32_t)__syncthreads_count(kept_old);
  if (tid == 0) chg[p] = nk > nold ? nk - nold : 0u;
}
template __global__ void k_merge_wide<512>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                           int, uint32_t*);
template __global__ void k_merge_wide<1024>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                            int, uint32_t*);
template __global__ void k_merge_wide<2048>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                            int, uint32_t*);

// ── extraction (nndescent.rs:977-1002) ──
// Rows hold k entries; the graph takes the first ko (ko <= k) of each. With a
// renumbering (perm[p] = original index of internal point p, inv its inverse),
// output row o is internal row inv[o] and neighbour indices map back through
// perm; otherwise perm and inv are NULL.
__global__ void k_row_counts(const uint64_t* __restrict__ G, uint32_t n, uint32_t k, uint32_t ko,
                             const uint32_t* __restrict__ inv, uint32_t* __restrict__ cnt) {
  // one thread per entry s < ko of output row i; a row's entries are a sorted
  // prefix followed by EMPTY, so the last non-empty one (or an empty first)
  // gives the count. n * ko < 2^32.
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ko) return;
  const uint32_t i = e / ko, s = e - i * ko;
  const uint64_t* r = G + (size_t)(inv ? inv[i] : i) * k;
  const bool here = r[s] != EMPTY_KEY;
  const bool next = s + 1 < ko && r[s + 1] != EMPTY_KEY;
  if (s == 0 && !here) cnt[i] = 0;
  if (here && !next) cnt[i] = s + 1;
}
__global__ void k_extract(const uint64_t* __restrict__ G, uint32_t n, uint32_t k, uint32_t ko,
                          const uint32_t* __restrict__ perm, const uint32_t* __restrict__ inv,
                          const uint32_t* __restrict__ off, uint32_t* __restrict__ indices,
                          float* __restrict__ dists) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;  // n * ko < 2^32
  if (e >= n * ko) return;
  const uint32_t i = e / ko, s = e - i * ko;
  uint64_t key = G[(size_t)(inv ? inv[i] : i) * k + s];
  if (key == EMPTY_KEY) return;
  uint32_t o = off[i] + s;
  indices[o] = perm ? perm[key_idx(key)] : key_idx(key);
  // sqrt in FP64 then rounded to FP32: with 53 >= 2*24 + 2 bits the result is
  // the correctly rounded f32 sqrt (gfx950's f32 sqrt is not), as Rust's f32::sqrt.
  dists[o] = (float)__dsqrt_rn((double)key_dist(key));
}

// ── renumbering into a locality order ──
__global__ void k_perm_inverse(const uint32_t* __restrict__ perm, uint32_t n,
                               uint32_t* __restrict__ inv) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) inv[perm[p]] = p;
}
// Y[p] = X[perm[p]], rows of dp floats (dp a multiple of 4).
__global__ void k_perm_rows(const float* __restrict__ X, uint32_t dp, const uint32_t* __restrict__ perm,
                            uint32_t n, float* __restrict__ Y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c4 = dp / 4;
  if (e >= (size_t)n * c4) return;
  const uint32_t p = (uint32_t)(e / c4), c = (uint32_t)(e - (size_t)p * c4);
  reinterpret_cast<float4*>(Y)[e] = reinterpret_cast<const float4*>(X)[(size_t)perm[p] * c4 + c];
}
// ids[e] = inv[ids[e]]: point indices of the forest's leaf lists in the new
// numbering.
__global__ void k_renumber(uint32_t* __restrict__ ids, size_t total, const uint32_t* __restrict__ inv) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) ids[e] = inv[ids[e]];
}

// explicit instantiations used by the host code
template __global__ void k_join<0>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                   const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                                   uint32_t, JoinCells, const float*, uint64_t*, uint8_t*, uint32_t,
                                   uint64_t, uint64_t, uint32_t*);
template __global__ void k_join<50>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                    const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                                    uint32_t, JoinCells, const float*, uint64_t*, uint8_t*, uint32_t,
                                    uint64_t, uint64_t, uint32_t*);

#define PSQ_MERGE_INST(NR, NT)                                                                 \
  template __global__ void k_merge<NR, NT>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t,   \
                                           uint8_t*, int, uint32_t*);
PSQ_MERGE_INST(1, 2)
PSQ_MERGE_INST(2, 4)
PSQ_MERGE_INST(4, 8)

}  // namespace psq
