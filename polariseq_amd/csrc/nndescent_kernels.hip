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

// ── sampling (nndescent.rs:853-909) ──
__global__ void k_rev_count(const uint64_t* __restrict__ G, size_t nk, uint32_t* __restrict__ rcount) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nk) return;
  uint64_t key = G[e];
  if (key != EMPTY_KEY) atomicAdd(&rcount[key_idx(key)], 1u);
}

// Reverse proposals grouped by destination. The order inside a group depends on
// atomic arrival and is irrelevant: the consumer is a set function.
// Entry: (class << 63) | (prio24 << 32) | src, class 0 = new, 1 = old.
__global__ void k_rev_scatter(const uint64_t* __restrict__ G, uint32_t n, uint32_t k, uint64_t seed,
                              uint64_t iter, const uint32_t* __restrict__ roff,
                              uint32_t* __restrict__ rcount, uint64_t* __restrict__ rev) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * k) return;
  uint64_t key = G[e];
  if (key == EMPTY_KEY) return;
  uint32_t i = (uint32_t)(e / k);
  uint32_t j = key_idx(key);
  uint64_t cls = (key & 1) ? 0ull : 1ull;
  uint64_t pr = hash_prio24(seed, iter, e);
  uint32_t pos = roff[j] + atomicSub(&rcount[j], 1u) - 1u;
  rev[pos] = (cls << 63) | (pr << 32) | (uint64_t)i;
}

// One wave per row i. Candidates: the row's own entries (forward, priority of
// slot i*k+s) and the reverse entries from rev[roff[i]..roff[i+1]). Per flag
// class, a neighbour seen twice keeps its smaller priority; then the m smallest
// (priority, neighbour) keys form the new and old lists. Entries of the row
// that made the new list lose their new flag (nndescent.rs:894-909).
__global__ __launch_bounds__(256) void k_sample_select(
    uint64_t* __restrict__ G, uint32_t n, uint32_t k, uint32_t m, uint64_t seed, uint64_t iter,
    const uint32_t* __restrict__ roff, const uint64_t* __restrict__ rev,
    uint32_t* __restrict__ newl, uint32_t* __restrict__ oldl, uint32_t* __restrict__ ncnt,
    uint64_t* __restrict__ slots, uint32_t S) {
  __shared__ uint64_t s_fk[4][WAVE];
  __shared__ uint32_t s_kill[4][WAVE];
  const int w = threadIdx.x / WAVE;
  const int l = lane_id();
  const uint32_t i = blockIdx.x * 4 + w;
  const bool active = i < n;  // wave-uniform
  const uint64_t fkey = (active && l < (int)k) ? G[(size_t)i * k + l] : EMPTY_KEY;
  const bool fvalid = fkey != EMPTY_KEY;
  uint64_t fk = EMPTY_KEY;
  if (fvalid) {
    uint64_t cls = (fkey & 1) ? 0ull : 1ull;
    uint64_t pr = hash_prio24(seed, iter, (uint64_t)i * k + l);
    fk = (cls << 63) | (pr << 32) | key_idx(fkey);
  }
  s_fk[w][l] = fk;
  s_kill[w][l] = 0;
  __syncthreads();
  const uint32_t r0 = active ? roff[i] : 0u, r1 = active ? roff[i + 1] : 0u;
  // pass 1: a reverse entry with a smaller priority than the matching forward
  // entry of the same class removes the forward one.
  for (uint32_t base = r0; base < r1; base += WAVE) {
    uint32_t e = base + l;
    if (e < r1) {
      uint64_t rv = rev[e];
      uint32_t src = (uint32_t)rv;
      uint64_t cp = rv >> 32;  // class and priority
      for (uint32_t f = 0; f < k; ++f) {
        uint64_t fv = s_fk[w][f];
        if (fv != EMPTY_KEY && (uint32_t)fv == src && (fv >> 63) == (cp >> 31)) {
          if ((cp & 0x7FFFFFFFull) < ((fv >> 32) & 0x7FFFFFFFull)) s_kill[w][f] = 1;
        }
      }
    }
  }
  __syncthreads();
  // pass 2: selection, keeping the 64 smallest keys per class.
  uint64_t bestN, bestO;
  {
    uint64_t key = (fvalid && !s_kill[w][l]) ? fk : EMPTY_KEY;
    uint64_t kn = (key != EMPTY_KEY && (key >> 63) == 0) ? key : EMPTY_KEY;
    uint64_t ko = (key != EMPTY_KEY && (key >> 63) == 1) ? key : EMPTY_KEY;
    bestN = wave_sort64(kn);
    bestO = wave_sort64(ko);
  }
  for (uint32_t base = r0; base < r1; base += WAVE) {
    uint32_t e = base + l;
    uint64_t key = EMPTY_KEY;
    if (e < r1) {
      uint64_t rv = rev[e];
      uint32_t src = (uint32_t)rv;
      uint64_t cp = rv >> 32;
      bool dead = false;
      for (uint32_t f = 0; f < k; ++f) {
        uint64_t fv = s_fk[w][f];
        if (fv != EMPTY_KEY && (uint32_t)fv == src && (fv >> 63) == (cp >> 31)) {
          if (((fv >> 32) & 0x7FFFFFFFull) <= (cp & 0x7FFFFFFFull)) dead = true;
        }
      }
      if (!dead) key = rv;
    }
    uint64_t kn = (key != EMPTY_KEY && (key >> 63) == 0) ? key : EMPTY_KEY;
    uint64_t ko = (key != EMPTY_KEY && (key >> 63) == 1) ? key : EMPTY_KEY;
    bestN = wave_merge_keep64(bestN, wave_sort64(kn));
    bestO = wave_merge_keep64(bestO, wave_sort64(ko));
  }
  const bool inN = (l < (int)m) && bestN != EMPTY_KEY;
  const bool inO = (l < (int)m) && bestO != EMPTY_KEY;
  const uint32_t nn = __popcll(__ballot(inN));
  const uint32_t no = __popcll(__ballot(inO));
  // flag update: a wave-uniform loop over the new list (all lanes shuffle)
  const uint32_t me = key_idx(fkey);
  bool found = false;
  for (uint32_t r = 0; r < nn; ++r) {
    uint32_t v = (uint32_t)shfl64(bestN, (int)r);
    found |= (v == me);
  }
  if (active) {
    if (l < (int)m) {
      newl[(size_t)i * m + l] = inN ? (uint32_t)bestN : EMPTY;
      oldl[(size_t)i * m + l] = inO ? (uint32_t)bestO : EMPTY;
    }
    if (l == 0) ncnt[i] = nn | (no << 16);
    if (fvalid && (fkey & 1) && found) G[(size_t)i * k + l] = fkey & ~1ull;
    // Bucket layout for this iteration's join: the row's neighbours pinned,
    // ascending by index inside each bucket, at most BUCKET per bucket; the
    // other slots empty. Every row is rewritten, so no bucket state carries over.
    const uint32_t bmask = S / BUCKET - 1;
    const uint32_t mb = fvalid ? (slot_hash(seed, iter, me) & bmask) : EMPTY;
    uint32_t rank = 0;
    for (uint32_t s = 0; s < k; ++s) {
      const uint32_t bs = __shfl(mb, (int)s, WAVE), is = __shfl(me, (int)s, WAVE);
      rank += (fvalid && bs == mb && is < me) ? 1u : 0u;
    }
    for (uint32_t e0 = 0; e0 < S; e0 += WAVE) {
      const uint32_t e = e0 + l;
      const uint32_t eb = e / BUCKET, ej = e % BUCKET;
      uint64_t val = EMPTY_KEY;
      for (uint32_t s = 0; s < k; ++s) {
        const uint32_t bs = __shfl(mb, (int)s, WAVE), rs = __shfl(rank, (int)s, WAVE);
        const uint32_t is = __shfl(me, (int)s, WAVE);
        if (bs == eb && rs == ej) val = pin_key(is);
      }
      if (e < S) slots[(size_t)i * S + e] = val;
    }
  }
}

// ── local join (nndescent.rs:912-969) ──
// One wave (one block) per cell i. The candidate rows (new list, then old
// list) are staged in LDS with coalesced 16-byte loads, at a row stride RS
// (floats, RS / 4 odd) that spreads rows over the LDS banks. The cell's pairs,
// new x new (a < b) then new x old with different points, are numbered and
// dealt to the 64 lanes, so every lane computes a distance per round. A pair
// whose distance beats an endpoint's threshold (snapshot of that row's worst)
// proposes the other endpoint to it; proposals are batched through the wave's
// LDS list. D = 50 fixes the dimension at compile time; D = 0 reads it at run
// time. Both use the reference's operation order (simd.rs:45-62).
template <int D>
__device__ inline float lds_sq_dist(const float* pa, const float* pb, uint32_t d) {
  if (D > 0) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ch = 0; ch < D / 8; ++ch) {
      const float4 a0 = *reinterpret_cast<const float4*>(pa + ch * 8);
      const float4 a1 = *reinterpret_cast<const float4*>(pa + ch * 8 + 4);
      const float4 b0 = *reinterpret_cast<const float4*>(pb + ch * 8);
      const float4 b1 = *reinterpret_cast<const float4*>(pb + ch * 8 + 4);
      float t;
      t = a0.x - b0.x; acc[0] = acc[0] + t * t;
      t = a0.y - b0.y; acc[1] = acc[1] + t * t;
      t = a0.z - b0.z; acc[2] = acc[2] + t * t;
      t = a0.w - b0.w; acc[3] = acc[3] + t * t;
      t = a1.x - b1.x; acc[4] = acc[4] + t * t;
      t = a1.y - b1.y; acc[5] = acc[5] + t * t;
      t = a1.z - b1.z; acc[6] = acc[6] + t * t;
      t = a1.w - b1.w; acc[7] = acc[7] + t * t;
    }
    float tail = 0.f;
#pragma unroll
    for (int j = (D / 8) * 8; j < D; ++j) {
      const float t = pa[j] - pb[j];
      tail = tail + t * t;
    }
    return (((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]))) +
           tail;
  }
  return sq_dist_ref(pa, pb, (int)d);
}

template <int D>
__global__ __launch_bounds__(64) void k_join(
    const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp, uint32_t RS,
    const uint32_t* __restrict__ newl, const uint32_t* __restrict__ oldl,
    const uint32_t* __restrict__ ncnt, uint32_t m, const float* __restrict__ thr,
    uint64_t* __restrict__ slots, uint8_t* __restrict__ touched, uint32_t S, uint64_t seed,
    uint64_t iter, unsigned long long* __restrict__ counters) {
  // LDS: pb_key[PB_CAP] u64 | rows[2m][RS] f32 | ids[2m] | thr[2m] | pb_row[PB_CAP]
  extern __shared__ __attribute__((aligned(16))) float lds[];
  uint64_t* pb_key = reinterpret_cast<uint64_t*>(lds);
  float* rows = lds + 2 * PB_CAP;
  uint32_t* s_id = reinterpret_cast<uint32_t*>(rows + (size_t)2 * m * RS);
  float* s_thr = reinterpret_cast<float*>(s_id + 2 * m);
  uint32_t* pb_row = reinterpret_cast<uint32_t*>(s_thr + 2 * m);
  const uint32_t l = (uint32_t)lane_id();
  const uint32_t i = blockIdx.x;
  if (i >= n) return;  // block-uniform
  const uint32_t cnt = ncnt[i];
  const uint32_t nn = cnt & 0xFFFF, no = cnt >> 16, nc = nn + no;
  if (nn == 0) return;
  const uint32_t bmask = S / BUCKET - 1;
  for (uint32_t s = l; s < nc; s += WAVE) {
    const uint32_t v = s < nn ? newl[(size_t)i * m + s] : oldl[(size_t)i * m + (s - nn)];
    s_id[s] = v;
    s_thr[s] = thr[v];
  }
  __syncthreads();
  const uint32_t c4 = D > 0 ? (uint32_t)((D + 3) / 4) : (d + 3) / 4;
  for (uint32_t e = l; e < nc * c4; e += WAVE) {
    const uint32_t s = e / c4, c = e - s * c4;
    const float4 v = *reinterpret_cast<const float4*>(Xp + (size_t)s_id[s] * dp + 4 * c);
    *reinterpret_cast<float4*>(rows + (size_t)s * RS + 4 * c) = v;
  }
  __syncthreads();
  const uint32_t P1 = nn * (nn - 1) / 2, P = P1 + nn * no;
  uint32_t pcnt = 0;
  uint64_t pairs = 0;
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
        dist = lds_sq_dist<D>(rows + (size_t)a * RS, rows + (size_t)b * RS, d);
        pa = dist < s_thr[a];
        pb = dist < s_thr[b];
      }
    }
    pb_append(pb_row, pb_key, pcnt, pa, ia, make_key(dist, ib, 1));
    pb_append(pb_row, pb_key, pcnt, pb, ib, make_key(dist, ia, 1));
    pb_flush(pb_row, pb_key, pcnt, false, slots, touched, S, bmask, seed, iter);
  }
  pb_flush(pb_row, pb_key, pcnt, true, slots, touched, S, bmask, seed, iter);
  // counters[0] += pairs, counters[1] += nn + no (candidate rows gathered)
  uint64_t tot = pairs;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += shfl_xor64(tot, o);
  if (l == 0) {
    ctr_add(counters, 0, (unsigned long long)tot);
    ctr_add(counters, 1, (unsigned long long)nc);
  }
}

// ── merge (the apply of nndescent.rs:307-324 as a set function) ──
// One wave per touched row: top-k of (row ∪ proposals), deduplicated on (dist,
// idx) keeping the smaller flag (an entry already in the row keeps its flag).
// Pinned slots are skipped. Word 2 of the striped counters `changed` counts
// entries of the new row absent from the old.
template <int NR>
__global__ __launch_bounds__(256) void k_merge(uint64_t* __restrict__ G, uint32_t n, uint32_t k,
                                               uint64_t* __restrict__ slots, uint32_t S,
                                               uint8_t* __restrict__ touched, float* __restrict__ thr,
                                               int force, unsigned long long* __restrict__ changed) {
  const int w = threadIdx.x / WAVE;
  const int l = lane_id();
  const uint32_t p = blockIdx.x * 4 + w;
  if (p >= n) return;
  if (!force && !touched[p]) return;
  uint64_t* row = G + (size_t)p * k;
  uint64_t* sl = slots + (size_t)p * S;
  const uint64_t orow = l < (int)k ? row[l] : EMPTY_KEY;
  uint64_t v[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const uint32_t e = r * WAVE + l;
    if (e < S) {
      v[r] = bucket_to_row(sl[e]);
    } else if (e < S + k) {
      v[r] = row[e - S];
    } else {
      v[r] = EMPTY_KEY;
    }
  }
  if (l == 0) touched[p] = 0;
  wave_sort_n<NR>(v);
  // dedup on (dist, idx): drop an element equal to its predecessor's
  uint32_t before = 0;
  uint32_t nkept = 0;
  uint32_t nchanged = 0;
  uint64_t lastkey = EMPTY_KEY;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    uint64_t prev = shfl64(v[r], (l + WAVE - 1) & (WAVE - 1));
    if (l == 0) prev = (r > 0) ? lastkey : EMPTY_KEY;
    const bool has_prev = (r > 0) || (l > 0);
    const bool keep = v[r] != EMPTY_KEY && !(has_prev && (v[r] >> 1) == (prev >> 1));
    const uint64_t mask = __ballot(keep);
    const uint32_t rank = before + __popcll(mask & ((1ull << l) - 1ull));
    bool found = false;
    for (uint32_t s2 = 0; s2 < k; ++s2) {
      uint64_t o = shfl64(orow, (int)s2);
      found |= (o != EMPTY_KEY && (o >> 1) == (v[r] >> 1));
    }
    if (keep && rank < k) {
      row[rank] = v[r];
      nchanged += !found;
      if (rank == k - 1) thr[p] = key_dist(v[r]);
    }
    before += __popcll(mask);
    lastkey = shfl64(v[r], WAVE - 1);
  }
  nkept = before < k ? before : k;
  for (uint32_t pos = nkept + l; pos < k; pos += WAVE) row[pos] = EMPTY_KEY;
  if (nkept < k && l == 0) thr[p] = __builtin_inff();
  uint32_t tot = nchanged;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, WAVE);
  if (l == 0 && tot) ctr_add(changed, 2, (unsigned long long)tot);
}

// ── extraction (nndescent.rs:977-1002) ──
// Rows hold k entries; the graph takes the first ko (ko <= k) of each.
__global__ void k_row_counts(const uint64_t* __restrict__ G, uint32_t n, uint32_t k, uint32_t ko,
                             uint32_t* __restrict__ cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t* r = G + (size_t)i * k;
  uint32_t c = 0;
  while (c < ko && r[c] != EMPTY_KEY) ++c;
  cnt[i] = c;
}
__global__ void k_extract(const uint64_t* __restrict__ G, uint32_t n, uint32_t k, uint32_t ko,
                          const uint32_t* __restrict__ off, uint32_t* __restrict__ indices,
                          float* __restrict__ dists) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * ko) return;
  uint32_t i = (uint32_t)(e / ko), s = (uint32_t)(e - (size_t)i * ko);
  uint64_t key = G[(size_t)i * k + s];
  if (key == EMPTY_KEY) return;
  uint32_t o = off[i] + s;
  indices[o] = key_idx(key);
  // sqrt in FP64 then rounded to FP32: with 53 >= 2*24 + 2 bits the result is
  // the correctly rounded f32 sqrt (gfx950's f32 sqrt is not), as Rust's f32::sqrt.
  dists[o] = (float)__dsqrt_rn((double)key_dist(key));
}

// explicit instantiations used by the host code
template __global__ void k_join<0>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                   const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                                   const float*, uint64_t*, uint8_t*, uint32_t, uint64_t, uint64_t,
                                   unsigned long long*);
template __global__ void k_join<50>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                    const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                                    const float*, uint64_t*, uint8_t*, uint32_t, uint64_t, uint64_t,
                                    unsigned long long*);
template __global__ void k_merge<2>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                    float*, int, unsigned long long*);
template __global__ void k_merge<4>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                    float*, int, unsigned long long*);
template __global__ void k_merge<8>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                    float*, int, unsigned long long*);

}  // namespace psq
