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
    uint32_t* __restrict__ newl, uint32_t* __restrict__ oldl, uint32_t* __restrict__ ncnt) {
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
  }
}

// ── local join (nndescent.rs:912-969) ──
// One wave per cell i. Lane t holds candidate slot t (and t + 64 when CPL = 2):
// slots [0, nn) are the new list, [nn, nn + no) the old list. The new rows are
// staged in LDS; for each new entry a (a wave-uniform loop), every lane computes
// sq_dist(new[a], its candidate) with the new row read by LDS broadcast and its
// own row in registers. Pairs: new×new with slot > a, and new×old with the
// candidate != new[a]. A pair proposes each endpoint to the other when its
// distance beats that endpoint's threshold (snapshot of the row's worst).
template <int D>
__device__ inline void load_row(const float* __restrict__ src, float (&q)[D]) {
  // rows are 128-byte aligned (padded stride)
  const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
  for (int t = 0; t < D / 4; ++t) {
    float4 v = s4[t];
    q[4 * t] = v.x;
    q[4 * t + 1] = v.y;
    q[4 * t + 2] = v.z;
    q[4 * t + 3] = v.w;
  }
#pragma unroll
  for (int t = (D / 4) * 4; t < D; ++t) q[t] = src[t];
}

template <int D, int CPL>
__global__ __launch_bounds__(256) void k_join(
    const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp,
    const uint32_t* __restrict__ newl, const uint32_t* __restrict__ oldl,
    const uint32_t* __restrict__ ncnt, uint32_t m, const float* __restrict__ thr,
    uint64_t* __restrict__ slots, uint8_t* __restrict__ touched, uint32_t S, uint64_t seed,
    uint64_t iter, unsigned long long* __restrict__ counters) {
  extern __shared__ float s_rows[];  // [4 waves][m][dp]
  const int w = threadIdx.x / WAVE;
  const int l = lane_id();
  const uint32_t i = blockIdx.x * 4 + w;
  const uint32_t cnt = i < n ? ncnt[i] : 0u;  // wave-uniform
  const uint32_t nn = cnt & 0xFFFF, no = cnt >> 16;
  float* prow = s_rows + (size_t)w * m * dp;
  const uint32_t bmask = S / BUCKET - 1;

  uint32_t cand[CPL];
  float cthr[CPL];
  constexpr int DR = D > 0 ? D : 1;
  float q[CPL][DR];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const uint32_t s = l + c * WAVE;
    uint32_t v = EMPTY;
    if (s < nn)
      v = newl[(size_t)i * m + s];
    else if (s < nn + no)
      v = oldl[(size_t)i * m + (s - nn)];
    cand[c] = v;
    cthr[c] = v != EMPTY ? thr[v] : 0.0f;
    if (D > 0) {
      if (v != EMPTY) {
        load_row<DR>(Xp + (size_t)v * dp, q[c]);
      } else {
#pragma unroll
        for (int t = 0; t < DR; ++t) q[c][t] = 0.0f;
      }
    }
  }
  // stage the new rows (slots < nn live in lanes < nn of group 0, nn <= m <= 64)
  for (uint32_t e = l; e < nn * dp; e += WAVE) {
    uint32_t a = e / dp, t = e - a * dp;
    uint32_t v = newl[(size_t)i * m + a];
    prow[e] = Xp[(size_t)v * dp + t];
  }
  __syncthreads();
  uint64_t pairs = 0;
  for (uint32_t a = 0; a < nn; ++a) {
    const uint32_t p = (uint32_t)__shfl((int)cand[0], (int)a, WAVE);
    const float thr_p = __shfl(cthr[0], (int)a, WAVE);
    const float* pr = prow + (size_t)a * dp;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const uint32_t s = l + c * WAVE;
      const bool valid = (s < nn) ? (s > a) : (s < nn + no && cand[c] != p);
      if (__ballot(valid) == 0) continue;
      if (!valid) continue;
      ++pairs;
      float dist;
      if (D > 0)
        dist = sq_dist_fixed<DR>(pr, q[c]);
      else
        dist = sq_dist_ref(pr, Xp + (size_t)cand[c] * dp, (int)d);
      if (dist < thr_p) propose(slots, touched, p, make_key(dist, cand[c], 1), S, bmask, seed, iter);
      if (dist < cthr[c]) propose(slots, touched, cand[c], make_key(dist, p, 1), S, bmask, seed, iter);
    }
  }
  // counters[0] += pairs, counters[1] += nn + no (candidate rows gathered)
  uint64_t tot = pairs;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += shfl_xor64(tot, o);
  if (l == 0 && nn > 0) {
    atomicAdd(&counters[0], (unsigned long long)tot);
    atomicAdd(&counters[1], (unsigned long long)(nn + no));
  }
}

// ── merge (the apply of nndescent.rs:307-324 as a set function) ──
// One wave per touched row: top-k of (row ∪ slots), deduplicated on (dist, idx)
// keeping the smaller flag (an entry already in the row keeps its flag). The
// slots are reset. `changed` counts entries of the new row absent from the old.
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
      v[r] = sl[e];
      sl[e] = EMPTY_KEY;
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
  if (l == 0 && tot) atomicAdd(&changed[0], (unsigned long long)tot);
}

// ── extraction (nndescent.rs:977-1002) ──
__global__ void k_row_counts(const uint64_t* __restrict__ G, uint32_t n, uint32_t k,
                             uint32_t* __restrict__ cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t* r = G + (size_t)i * k;
  uint32_t c = 0;
  while (c < k && r[c] != EMPTY_KEY) ++c;
  cnt[i] = c;
}
__global__ void k_extract(const uint64_t* __restrict__ G, uint32_t n, uint32_t k,
                          const uint32_t* __restrict__ off, uint32_t* __restrict__ indices,
                          float* __restrict__ dists) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * k) return;
  uint32_t i = (uint32_t)(e / k), s = (uint32_t)(e - (size_t)i * k);
  uint64_t key = G[e];
  if (key == EMPTY_KEY) return;
  uint32_t o = off[i] + s;
  indices[o] = key_idx(key);
  // sqrt in FP64 then rounded to FP32: with 53 >= 2*24 + 2 bits the result is
  // the correctly rounded f32 sqrt (gfx950's f32 sqrt is not), as Rust's f32::sqrt.
  dists[o] = (float)__dsqrt_rn((double)key_dist(key));
}

// explicit instantiations used by the host code
template __global__ void k_join<0, 1>(const float*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                                      const uint32_t*, const uint32_t*, uint32_t, const float*,
                                      uint64_t*, uint8_t*, uint32_t, uint64_t, uint64_t,
                                      unsigned long long*);
template __global__ void k_join<0, 2>(const float*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                                      const uint32_t*, const uint32_t*, uint32_t, const float*,
                                      uint64_t*, uint8_t*, uint32_t, uint64_t, uint64_t,
                                      unsigned long long*);
template __global__ void k_join<50, 1>(const float*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                                       const uint32_t*, const uint32_t*, uint32_t, const float*,
                                       uint64_t*, uint8_t*, uint32_t, uint64_t, uint64_t,
                                       unsigned long long*);
template __global__ void k_join<50, 2>(const float*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                                       const uint32_t*, const uint32_t*, uint32_t, const float*,
                                       uint64_t*, uint8_t*, uint32_t, uint64_t, uint64_t,
                                       unsigned long long*);
template __global__ void k_merge<2>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                    float*, int, unsigned long long*);
template __global__ void k_merge<4>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                    float*, int, unsigned long long*);
template __global__ void k_merge<8>(uint64_t*, uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*,
                                    float*, int, unsigned long long*);

}  // namespace psq
