// Precision form of the local join: BF16 MFMA distance tiles with FP32
// accumulation, then an FP32 re-rank of each row's best 2k proposals.
//
// Same iteration structure as the plain form (crates/polariseq-core/src/
// nndescent.rs:912-969), with two changes:
//  * join: per cell i the candidate rows are centred on x_i in FP32 (all of
//    them are near x_i, so the centred vectors are short), rounded to BF16 and
//    staged in LDS with coalesced loads. Proposals are batched through the
//    wave's LDS list (see pb_append / pb_flush). One v_mfma_f32_16x16x32_bf16 pass gives the Gram tile of
//    new rows x all candidates; d~(p,q) = |p~|^2 + |q~|^2 - 2 p~.q~, with the
//    norms summed in FP32 from the same BF16 values. Proposals carry d~.
//  * merge: d~ differs between cells, so proposals are deduplicated on the
//    neighbour index; the 2k smallest new proposals (k = output neighbours)
//    are re-ranked with the exact f32 sq_dist (simd.rs:45-62 order) and the
//    row becomes the top of (row ∪ re-ranked), all exact. Rows keep 2k
//    entries by default and the graph takes the first k. Output distances are
//    therefore the plain form's, bit for bit, for every edge.
#include "psq_kernels.hpp"

namespace psq {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BF_KD = 64;  // padded dimensions (d <= 64)
constexpr int BF_RS = 72;  // LDS row stride in bf16: 144 bytes, so 16 rows hit distinct banks

// One wave (one block) per cell. LDS: pb_key[PB_CAP] u64 | rows[R][BF_RS] bf16
// | ids[R] | thr[R] | nrm[R] | pb_row[PB_CAP], R = 2m rounded up to 16.
__global__ __launch_bounds__(64) void k_join_bf16(
    const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp,
    const uint32_t* __restrict__ newl, const uint32_t* __restrict__ oldl,
    const uint32_t* __restrict__ ncnt, uint32_t m, const float* __restrict__ thr, float slack,
    uint64_t* __restrict__ slots, uint8_t* __restrict__ touched, uint32_t S, uint64_t seed,
    uint64_t iter, unsigned long long* __restrict__ counters) {
  extern __shared__ __attribute__((aligned(16))) uint64_t lds64[];
  const uint32_t R = (2 * m + 15) / 16 * 16;
  uint64_t* pb_key = lds64;
  __bf16* s_v = reinterpret_cast<__bf16*>(lds64 + PB_CAP);
  uint32_t* s_id = reinterpret_cast<uint32_t*>(s_v + (size_t)R * BF_RS);
  float* s_thr = reinterpret_cast<float*>(s_id + R);
  float* s_nrm = s_thr + R;
  uint32_t* pb_row = reinterpret_cast<uint32_t*>(s_nrm + R);
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
    s_thr[s] = thr[v] * slack;
  }
  __syncthreads();
  // centred rows, rounded to BF16: 16 chunks of 4 dimensions per row
  const float* xi = Xp + (size_t)i * dp;
  for (uint32_t e = l; e < nc * (BF_KD / 4); e += WAVE) {
    const uint32_t s = e / (BF_KD / 4), c = e % (BF_KD / 4);
    float4 dv = make_float4(0.f, 0.f, 0.f, 0.f);
    if (4 * c < dp) {
      const float4 xv = *reinterpret_cast<const float4*>(Xp + (size_t)s_id[s] * dp + 4 * c);
      const float4 x0 = *reinterpret_cast<const float4*>(xi + 4 * c);
      dv = make_float4(xv.x - x0.x, xv.y - x0.y, xv.z - x0.z, xv.w - x0.w);
    }
    __bf16* dst = s_v + (size_t)s * BF_RS + 4 * c;
    dst[0] = (__bf16)dv.x;
    dst[1] = (__bf16)dv.y;
    dst[2] = (__bf16)dv.z;
    dst[3] = (__bf16)dv.w;
  }
  __syncthreads();
  for (uint32_t s = l; s < nc; s += WAVE) {
    float nrm = 0.0f;
    const __bf16* row = s_v + (size_t)s * BF_RS;
    for (int t = 0; t < BF_KD; ++t) {
      const float bf = (float)row[t];
      nrm = fmaf(bf, bf, nrm);
    }
    s_nrm[s] = nrm;
  }
  __syncthreads();
  uint32_t pcnt = 0;
  uint64_t pairs = 0;
  const uint32_t rtiles = (nn + 15) / 16, ctiles = (nc + 15) / 16;
  const uint32_t lr = l % 16, lk = (l / 16) * 8;
  for (uint32_t rt = 0; rt < rtiles; ++rt) {
    for (uint32_t ct = 0; ct < ctiles; ++ct) {
      // skip tiles entirely below the new x new diagonal
      if (ct * 16 + 15 < rt * 16 && ct * 16 + 15 < nn) continue;
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < BF_KD; kk += 32) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(s_v + (size_t)(rt * 16 + lr) * BF_RS + kk + lk);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(s_v + (size_t)(ct * 16 + lr) * BF_RS + kk + lk);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
      }
      const uint32_t c = ct * 16 + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t r = rt * 16 + (l / 16) * 4 + v;
        bool pa = false, pb = false;
        uint32_t p = 0, q = 0;
        float dist = 0.0f;
        if (r < nn && c < nc) {
          p = s_id[r];
          q = s_id[c];
          if ((c < nn) ? (c > r) : (q != p)) {
            ++pairs;
            dist = s_nrm[r] + s_nrm[c] - 2.0f * acc[v];
            dist = dist > 0.0f ? dist : 0.0f;
            pa = dist < s_thr[r];
            pb = dist < s_thr[c];
          }
        }
        pb_append(pb_row, pb_key, pcnt, pa, p, make_key(dist, q, 1));
        pb_append(pb_row, pb_key, pcnt, pb, q, make_key(dist, p, 1));
        pb_flush(pb_row, pb_key, pcnt, false, slots, touched, S, bmask, seed, iter);
      }
    }
  }
  pb_flush(pb_row, pb_key, pcnt, true, slots, touched, S, bmask, seed, iter);
  uint64_t tot = pairs;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += shfl_xor64(tot, o);
  if (l == 0) {
    ctr_add(counters, 0, (unsigned long long)tot);
    ctr_add(counters, 1, (unsigned long long)nc);
  }
}

// One wave per touched row. Slots hold approximate keys. Steps: drop proposals
// whose neighbour is already in the row; inside each bucket (a neighbour always
// hashes to the same bucket) keep a neighbour's smallest key; take the rr
// smallest survivors (rr = twice the output k), recompute their exact f32
// distance, and keep the top k of (row ∪ re-ranked), k being the row size.
// NR: S / 64 registers per lane for the slots.
template <int NR>
__global__ __launch_bounds__(256) void k_merge_rerank(
    const float* __restrict__ Xp, uint32_t d, uint32_t dp, uint64_t* __restrict__ G, uint32_t n,
    uint32_t k, uint32_t rr, uint64_t* __restrict__ slots, uint32_t S, uint8_t* __restrict__ touched,
    float* __restrict__ thr, unsigned long long* __restrict__ changed) {
  const int w = threadIdx.x / WAVE, l = lane_id();
  const uint32_t p = blockIdx.x * 4 + w;
  if (p >= n) return;
  if (!touched[p]) return;
  uint64_t* row = G + (size_t)p * k;
  uint64_t* sl = slots + (size_t)p * S;
  const uint64_t orow = l < (int)k ? row[l] : EMPTY_KEY;
  uint64_t v[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const uint32_t e = r * WAVE + l;
    v[r] = e < S ? bucket_to_row(sl[e]) : EMPTY_KEY;
  }
  if (l == 0) touched[p] = 0;
  // 1. neighbours already in the row
  for (uint32_t s = 0; s < k; ++s) {
    const uint64_t o = shfl64(orow, (int)s);
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (v[r] != EMPTY_KEY && o != EMPTY_KEY && key_idx(o) == key_idx(v[r])) v[r] = EMPTY_KEY;
  }
  // 2. duplicates inside a bucket (8 consecutive lanes, sorted ascending):
  //    a later slot with the same neighbour is dropped.
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    bool dup = false;
#pragma unroll
    for (int j = 1; j < BUCKET; ++j) {
      const int src = (l & ~(BUCKET - 1)) | ((l - j) & (BUCKET - 1));
      const uint64_t o = shfl64(v[r], src);
      const bool earlier = (l & (BUCKET - 1)) >= j;
      dup |= earlier && v[r] != EMPTY_KEY && o != EMPTY_KEY && key_idx(o) == key_idx(v[r]);
    }
    if (dup) v[r] = EMPTY_KEY;
  }
  // 3. the rr smallest approximate keys
  wave_sort_n<NR>(v);
  // ranks are positions in the sorted order: element e = r * 64 + lane
  // 4. exact distances for ranks < rr (rr <= 128)
  const float* xp = Xp + (size_t)p * dp;
  uint64_t ex[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const uint32_t e = r * WAVE + l;
    ex[r] = EMPTY_KEY;
    if (e < rr && v[r] != EMPTY_KEY) {
      const uint32_t q = key_idx(v[r]);
      const float dd = sq_dist_ref(xp, Xp + (size_t)q * dp, (int)d);
      ex[r] = make_key(dd, q, 1);
    }
  }
  // 5. top k of (row ∪ re-ranked): row entries and the re-ranked keys
  //    (k + rr <= 3 * 64 elements) sorted together.
  uint64_t u[4];
  u[0] = orow;
#pragma unroll
  for (int r = 0; r < 3; ++r) u[r + 1] = (r < NR) ? ex[r] : EMPTY_KEY;
  wave_sort_n<4>(u);
  uint32_t before = 0, nchanged = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const bool keep = u[r] != EMPTY_KEY;
    const uint64_t mask = __ballot(keep);
    const uint32_t rank = before + __popcll(mask & ((1ull << l) - 1ull));
    bool found = false;
    for (uint32_t s = 0; s < k; ++s) {
      const uint64_t o = shfl64(orow, (int)s);
      found |= (o != EMPTY_KEY && o == u[r]);
    }
    if (keep && rank < k) {
      row[rank] = u[r];
      nchanged += !found;
      if (rank == k - 1) thr[p] = key_dist(u[r]);
    }
    before += __popcll(mask);
  }
  const uint32_t nkept = before < k ? before : k;
  for (uint32_t pos = nkept + l; pos < k; pos += WAVE) row[pos] = EMPTY_KEY;
  if (nkept < k && l == 0) thr[p] = __builtin_inff();
  uint32_t tot = nchanged;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, WAVE);
  if (l == 0 && tot) ctr_add(changed, 2, (unsigned long long)tot);
}

template __global__ void k_merge_rerank<1>(const float*, uint32_t, uint32_t, uint64_t*, uint32_t,
                                           uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*, float*,
                                           unsigned long long*);
template __global__ void k_merge_rerank<4>(const float*, uint32_t, uint32_t, uint64_t*, uint32_t,
                                           uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*, float*,
                                           unsigned long long*);
template __global__ void k_merge_rerank<2>(const float*, uint32_t, uint32_t, uint64_t*, uint32_t,
                                           uint32_t, uint32_t, uint64_t*, uint32_t, uint8_t*, float*,
                                           unsigned long long*);

}  // namespace psq
