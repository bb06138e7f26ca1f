// Device-side building blocks shared by the kNN kernels (gfx950, wave64).
//
// Hashes restate crates/polariseq-core/src/nndescent.rs:80-100; the 8-lane
// squared distance restates crates/polariseq-core/src/simd.rs:45-81 operation for
// operation (this file is compiled with -ffp-contract=off, as Rust never fuses).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psq {

constexpr uint32_t EMPTY = 0xFFFFFFFFu;
constexpr uint64_t EMPTY_KEY = ~0ull;
constexpr uint64_t GOLD = 0x9E3779B97F4A7C15ull;
constexpr int WAVE = 64;
constexpr int BUCKET = 8;  // slots per proposal bucket

__host__ __device__ inline uint64_t splitmix(uint64_t x) {
  x += GOLD;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// nndescent.rs:91-95, the 24-bit numerator of hash_unit.
__host__ __device__ inline uint32_t hash_prio24(uint64_t seed, uint64_t a, uint64_t b) {
  return (uint32_t)(splitmix(seed ^ splitmix(a) ^ splitmix(b * GOLD)) >> 40);
}
// nndescent.rs:97-100.
__host__ __device__ inline uint64_t hash_index(uint64_t seed, uint64_t a, uint64_t b, uint64_t n) {
  return splitmix(seed ^ splitmix(a) ^ splitmix(b + 0x12345678ull)) % n;
}
// Forest and bucket hashes (splitmix chains of the same finaliser).
__host__ __device__ inline uint64_t tree_hash(uint64_t tseed, uint64_t level, uint64_t a,
                                              uint64_t salt) {
  return splitmix(tseed ^ splitmix(level * GOLD + salt) ^ splitmix(a * GOLD + 0x51ED27ull));
}
// Proposal buckets: neighbour q goes to bucket slot_bucket(salt, q, bits) of a
// row's 2^bits buckets in an iteration, the top bits of a multiplicative hash
// of q xor a per-iteration salt (one multiply per key; the salt is uniform).
struct SlotHash {
  uint32_t salt, bits;
};
__host__ __device__ inline SlotHash slot_hash_of(uint64_t seed, uint64_t iter, uint32_t S) {
  uint32_t bits = 0;
  while ((8u << bits) < S) ++bits;  // S / 8 buckets
  return SlotHash{(uint32_t)(splitmix((seed ^ 0xA5A5A5A5DEADBEEFull) + (iter << 32)) >> 32), bits};
}
__host__ __device__ inline uint32_t slot_bucket(const SlotHash& h, uint32_t q) {
  return h.bits ? ((q ^ h.salt) * 0x9E3779B1u) >> (32 - h.bits) : 0u;
}

// Row key: (dist_bits << 32) | (idx << 1) | is_new. dist >= 0, so the unsigned
// order of keys is (dist, idx, flag); EMPTY_KEY sorts last.
__device__ inline uint64_t make_key(float d, uint32_t idx, uint32_t is_new) {
  return ((uint64_t)__float_as_uint(d) << 32) | ((uint64_t)idx << 1) | is_new;
}
__device__ inline uint32_t key_idx(uint64_t key) { return ((uint32_t)key) >> 1; }
__device__ inline float key_dist(uint64_t key) { return __uint_as_float((uint32_t)(key >> 32)); }

// ── wave primitives (64 lanes) ──
// Lane l receives lane l ^ m. For the masks of the sorting networks (constant
// after unrolling) the exchange avoids the LDS crossbar's addressed permute:
// xor 1 and 2 are DPP quad permutations, xor 8 a DPP rotation by 8 within each
// row of 16 lanes, xor 4 and 16 ds_swizzle in bit mode (and 0x1f, xor m) within
// each half of 32 lanes; other masks use ds_bpermute.
__device__ inline uint32_t xor_lane32(uint32_t v, int m) {
  const int x = (int)v;
  switch (m) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
    case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle(x, 0x101F);
    case 8: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);
    case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle(x, 0x401F);
    default: return (uint32_t)__shfl_xor(x, m, WAVE);
  }
}
__device__ inline uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = xor_lane32((uint32_t)v, m), hi = xor_lane32((uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t shfl64(uint64_t v, int src) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl(lo, src, WAVE);
  hi = __shfl(hi, src, WAVE);
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

// Block b of a one-block-per-item launch. A dispatch counts fewer than 2^32
// work-items per grid dimension, so launches of many blocks are two-dimensional
// (the host's block_grid): b = blockIdx.y * gridDim.x + blockIdx.x, with
// gridDim.x a multiple of XCDS whenever gridDim.y > 1. Blocks past the item
// count do nothing.
__device__ inline uint32_t block_id() { return blockIdx.y * gridDim.x + blockIdx.x; }
__device__ inline uint32_t grid_blocks() { return gridDim.y * gridDim.x; }

// Row of block b for the one-row-per-block kernels whose rows share data with
// nearby rows (the candidate rows of a join). The hardware deals blocks to the
// XCDS dies round-robin in the order of b and each die has its own L2, so XCD x
// takes the contiguous rows [x * c, (x + 1) * c), c = grid_blocks() / XCDS, in
// order. The grid holds a multiple of XCDS blocks; rows >= n do nothing.
constexpr uint32_t XCDS = 8;
__device__ inline uint32_t xcd_row() {
  const uint32_t b = block_id();
  return (b % XCDS) * (grid_blocks() / XCDS) + b / XCDS;
}
__host__ inline uint32_t xcd_grid(uint64_t n) { return (uint32_t)((n + XCDS - 1) / XCDS * XCDS); }

// Bitonic sort of one u64 per lane within each segment of W lanes (W = 16,
// 32, 64), each segment ascending by lane.
template <int W>
__device__ inline uint64_t wave_sort_seg(uint64_t v) {
  const int l = lane_id() & (W - 1);
#pragma unroll
  for (int size = 2; size <= W; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      uint64_t o = shfl_xor64(v, stride);
      bool up = ((l & size) == 0);
      bool lower = ((l & stride) == 0);
      uint64_t mn = v < o ? v : o, mx = v < o ? o : v;
      v = (up == lower) ? mn : mx;
    }
  }
  return v;
}
// Bitonic sort of one u64 per lane, ascending by lane.
__device__ inline uint64_t wave_sort64(uint64_t v) { return wave_sort_seg<WAVE>(v); }
// Two independent bitonic sorts of 32: lanes 0-31 and lanes 32-63, each
// ascending by lane.
__device__ inline uint64_t wave_sort32x2(uint64_t v) { return wave_sort_seg<WAVE / 2>(v); }
// Sort of a vector whose first `cnt` lanes hold keys and whose other lanes hold
// EMPTY_KEY (the largest key): the lanes past the first 16 or 32 are already in
// order, so a segment sort of 16 or 32 sorts the whole vector when cnt fits.
__device__ inline uint64_t wave_sort_prefix(uint64_t v, uint32_t cnt) {
  if (cnt <= 16) return wave_sort_seg<16>(v);
  if (cnt <= 32) return wave_sort_seg<32>(v);
  return wave_sort64(v);
}
// v holds an ascending-by-lane bitonic sequence's first half property: sort a
// bitonic sequence of 64 ascending.
__device__ inline uint64_t wave_bitonic_merge64(uint64_t v) {
  const int l = lane_id();
#pragma unroll
  for (int stride = WAVE >> 1; stride > 0; stride >>= 1) {
    uint64_t o = shfl_xor64(v, stride);
    bool lower = ((l & stride) == 0);
    uint64_t mn = v < o ? v : o, mx = v < o ? o : v;
    v = lower ? mn : mx;
  }
  return v;
}
// best, chunk: ascending by lane. Returns the 64 smallest of their union, ascending.
__device__ inline uint64_t wave_merge_keep64(uint64_t best, uint64_t chunk_sorted) {
  uint64_t rev = shfl64(chunk_sorted, WAVE - 1 - lane_id());
  uint64_t m = best < rev ? best : rev;
  return wave_bitonic_merge64(m);
}

// Two u64 per lane: element e = r * 64 + lane. Ascending bitonic sort of 128.
__device__ inline void wave_sort128(uint64_t& a, uint64_t& b) {
  const int l = lane_id();
#pragma unroll
  for (int size = 2; size <= 2 * WAVE; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride == WAVE) {
        // pairs (a at e=l, b at e=64+l); direction from bit `size` of e (size >= 128 -> up)
        bool up = true;
        uint64_t mn = a < b ? a : b, mx = a < b ? b : a;
        a = up ? mn : mx;
        b = up ? mx : mn;
      } else {
        uint64_t oa = shfl_xor64(a, stride), ob = shfl_xor64(b, stride);
        bool lower = ((l & stride) == 0);
        bool upa = ((l & size) == 0);
        bool upb = (((WAVE + l) & size) == 0);
        uint64_t mna = a < oa ? a : oa, mxa = a < oa ? oa : a;
        uint64_t mnb = b < ob ? b : ob, mxb = b < ob ? ob : b;
        a = (upa == lower) ? mna : mxa;
        b = (upb == lower) ? mnb : mxb;
      }
    }
  }
}

// ── distances ──
// crates/polariseq-core/src/simd.rs:45-62 for a row `p` (any address space)
// against row `q`, both of length d: 8 lane sums over chunks of 8, the fixed
// reduce ((a0+a4)+(a1+a5))+((a2+a6)+(a3+a7)), then the sequential tail.
template <typename PA, typename PB>
__device__ inline float sq_dist_ref(PA p, PB q, int d) {
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int c = d >> 3;
  for (int ch = 0; ch < c; ++ch) {
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      float t = p[ch * 8 + l] - q[ch * 8 + l];
      acc[l] = acc[l] + t * t;
    }
  }
  float tail = 0.f;
  for (int i = c * 8; i < d; ++i) {
    float t = p[i] - q[i];
    tail = tail + t * t;
  }
  return (((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]))) + tail;
}
// crates/polariseq-core/src/simd.rs:22-37.
template <typename PA, typename PB>
__device__ inline float dot_ref(PA p, PB q, int d) {
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int c = d >> 3;
  for (int ch = 0; ch < c; ++ch) {
#pragma unroll
    for (int l = 0; l < 8; ++l) acc[l] = acc[l] + p[ch * 8 + l] * q[ch * 8 + l];
  }
  float tail = 0.f;
  for (int i = c * 8; i < d; ++i) tail = tail + p[i] * q[i];
  return (((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]))) + tail;
}

// NR u64 per lane, element e = r * 64 + lane. Ascending bitonic sort of 64*NR keys.
template <int NR>
__device__ inline void wave_sort_n(uint64_t (&v)[NR]) {
  const int l = lane_id();
#pragma unroll
  for (int size = 2; size <= WAVE * NR; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= WAVE) {
        const int rs = stride / WAVE;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int rp = r ^ rs;
          if (r < rp) {
            const bool up = (((r * WAVE + l) & size) == 0);
            uint64_t x = v[r], y = v[rp];
            uint64_t mn = x < y ? x : y, mx = x < y ? y : x;
            v[r] = up ? mn : mx;
            v[rp] = up ? mx : mn;
          }
        }
      } else {
        const bool lower = ((l & stride) == 0);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          uint64_t o = shfl_xor64(v[r], stride);
          const bool up = (((r * WAVE + l) & size) == 0);
          uint64_t mn = v[r] < o ? v[r] : o, mx = v[r] < o ? o : v[r];
          v[r] = (up == lower) ? mn : mx;
        }
      }
    }
  }
}

// simd.rs:45-62 with the dimension fixed at compile time, so a register array
// can hold one operand. Same operation order as sq_dist_ref.
template <int D, typename PA, typename PB>
__device__ inline float sq_dist_fixed(PA p, PB q) {
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ch = 0; ch < D / 8; ++ch) {
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      float t = p[ch * 8 + l] - q[ch * 8 + l];
      acc[l] = acc[l] + t * t;
    }
  }
  float tail = 0.f;
#pragma unroll
  for (int i = (D / 8) * 8; i < D; ++i) {
    float t = p[i] - q[i];
    tail = tail + t * t;
  }
  return (((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]))) + tail;
}

// ── statistics counters ──
// Counters are striped over CTR_STRIPES lines of CTR_WORDS u64: a block adds to
// its own stripe and k_ctr_reduce sums the stripes. Words: 0 pairs, 1 candidate
// rows, 2 changed entries, 3 exact pairs, 4 proposals.
constexpr uint32_t CTR_STRIPES = 256, CTR_WORDS = 8;
__device__ inline void ctr_add(unsigned long long* ctr, uint32_t word, unsigned long long v) {
  atomicAdd(&ctr[(size_t)(blockIdx.x % CTR_STRIPES) * CTR_WORDS + word], v);
}

// The join's per-cell record: JOIN_REC_WORDS u32 per cell, word w counting what
// counter word w counts (0 pairs, 1 candidate rows, 3 exact pairs, 4 proposals;
// 2, 5, 6 and 7 stay zero). The launch over every cell writes every cell's
// record, zeros for the cells it leaves to a deferred launch, which runs later
// on the same stream and overwrites them; k_join_rec_reduce then adds the
// records to the counter stripes. One 32-byte store per cell replaces four
// atomics per cell on CTR_STRIPES shared lines.
constexpr uint32_t JOIN_REC_WORDS = 8;
__device__ inline void join_record(uint32_t* rec, uint32_t i, uint32_t l, uint32_t pairs,
                                   uint32_t rows, uint32_t exact, uint32_t sent) {
  if (l < JOIN_REC_WORDS) {
    const uint32_t v = l == 0 ? pairs : l == 1 ? rows : l == 3 ? exact : l == 4 ? sent : 0u;
    rec[(size_t)i * JOIN_REC_WORDS + l] = v;
  }
}

// ── proposals: bucket cascade ──
// Each row owns S/8 buckets of 8 slots; a neighbour q goes to bucket
// slot_bucket(slot_hash_of(seed, iter, S), q). Before the join, a row's buckets are
// laid out with its current neighbours "pinned": pinned key = idx << 1 (dist
// field 0), ascending, at most 8 per bucket. A proposal is stored as its row
// key plus 1 << 32 (dist field + 1), so it sorts after every pinned key and
// pinned slots never move. A proposal of a neighbour the row already holds is
// rejected by reading the bucket, with no atomic.
__device__ inline uint64_t pin_key(uint32_t idx) { return (uint64_t)idx << 1; }
__device__ inline uint64_t bucket_key(uint64_t row_key) { return row_key + (1ull << 32); }
// Bucket slot -> row key; EMPTY_KEY for empty and pinned slots.
__device__ inline uint64_t bucket_to_row(uint64_t b) {
  return (b == EMPTY_KEY || (b >> 32) == 0) ? EMPTY_KEY : b - (1ull << 32);
}

// Insert a proposal (row key `key`, distance finite) into its bucket. A bucket
// holds the smallest distinct keys offered to it (pinned keys included), in no
// particular order: a key that is absent and below the bucket's largest slot
// replaces that slot with one compare-and-swap (the first such slot when
// several are equal, e.g. empty). A failed swap returns the slot's current
// value and the choice is made again. Slots only ever decrease, so every value
// read is the slot's value at some earlier time; two inserts of one key cannot
// both succeed in different slots (each would have to see the other's slot as
// the larger, older one), and the content after all proposals is the set of
// the 8 smallest distinct keys, whatever the order of arrival.
// "Absent" is decided on the low word: every key of neighbour q in a row's
// buckets, pinned (q << 1) or proposed (row key + 2^32), has low word
// (q << 1) | flag, and a proposal of q always carries the same distance (the
// distance of one pair is computed one way), so a low-word match is a match of
// the whole key. An empty slot's low word (2^32 - 1) matches no index < 2^31 - 1.
// Among several slots holding the largest value (empty slots: keys are
// distinct), a key takes the first in its own rotation of the 8 slots (by a
// hash of q), so concurrent inserts of different keys spread over the empty
// slots instead of all swapping on the first one (27% of the swaps of the first
// BF16 join failed that way). Two inserts of one key still cannot both succeed:
// they scan in the same order, the empty slots only become fewer, so the later
// reader either picks the same slot or finds the earlier pick already filled
// (and its own swap on an older view fails).
__device__ inline void propose(uint64_t* __restrict__ slots, uint8_t* __restrict__ touched,
                               uint32_t row, uint64_t key, uint32_t S, const SlotHash& sh) {
  const uint32_t q = key_idx(key);
  const uint32_t b = slot_bucket(sh, q);
  uint64_t* bk = slots + (size_t)row * S + (size_t)b * BUCKET;
  const ulonglong2* b2 = reinterpret_cast<const ulonglong2*>(bk);
  const uint64_t bkey = bucket_key(key);
  const uint32_t qq = q << 1;
  const uint32_t rot = (q * 0x85EBCA6Bu) >> 29;  // 0..7
  const ulonglong2 a = b2[0], bb = b2[1], c = b2[2], e = b2[3];
  uint64_t v[8] = {a.x, a.y, bb.x, bb.y, c.x, c.y, e.x, e.y};
  for (;;) {
    bool dup = false;
    uint64_t mx = v[0];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      dup |= ((uint32_t)v[j] & ~1u) == qq;
      mx = v[j] > mx ? v[j] : mx;
    }
    if (dup || bkey >= mx) return;  // present, or 8 smaller keys present
    uint32_t eq = 0;  // slots holding the largest value
#pragma unroll
    for (int j = 0; j < 8; ++j) eq |= (v[j] == mx ? 1u : 0u) << j;
    const uint32_t rq = ((eq >> rot) | (eq << (8 - rot))) & 0xFFu;  // not 0: mx is a slot's value
    const int jm = (int)((__builtin_ctz(rq) + rot) & 7u);
    const uint64_t old = atomicCAS((unsigned long long*)&bk[jm], (unsigned long long)mx,
                                   (unsigned long long)bkey);
    if (old == mx) break;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = j == jm ? old : v[j];
  }
  touched[row] = 1;
}

// sq_dist of two rows in LDS (16-byte aligned): D = 50 fixed at compile time with
// vector reads, D = 0 any d. Operation order of simd.rs:45-62.
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

// ── join helpers: pair enumeration and the per-wave proposal list ──
// Pairs of one cell, numbered t = 0 .. P1 + nn*no - 1: first the new x new
// upper triangle (a < b < nn) in row order, then new x old (a < nn, b = nn + j).
// off(a) = a * (2nn - a - 1) / 2 pairs precede row a of the triangle.
__host__ __device__ inline void pair_decode(uint32_t t, uint32_t nn, uint32_t no, uint32_t P1,
                                            uint32_t& a, uint32_t& b) {
  if (t < P1) {
    const float fn = 2.0f * (float)nn - 1.0f;
    const float disc = fn * fn - 8.0f * (float)t;
    int32_t aa = (int32_t)((fn - sqrtf(disc > 0.0f ? disc : 0.0f)) * 0.5f);
    const int32_t amax = (int32_t)nn - 2;
    aa = aa < 0 ? 0 : (aa > amax ? amax : aa);
    const uint32_t n2 = 2 * nn - 1;
    while (aa > 0 && (uint32_t)aa * (n2 - (uint32_t)aa) / 2 > t) --aa;
    while (aa < amax && (uint32_t)(aa + 1) * (n2 - (uint32_t)(aa + 1)) / 2 <= t) ++aa;
    a = (uint32_t)aa;
    b = a + 1 + (t - a * (n2 - a) / 2);
  } else {
    const uint32_t u = t - P1;
    a = u / no;
    b = nn + (u - a * no);
  }
}

// Proposals a wave finds are appended to an LDS list and sent 64 at a time,
// one per lane, so a batch's bucket reads and atomics are in flight together.
// Orders one wave's LDS accesses when the other waves of its block are not
// synchronised with it: a wave executes its LDS operations in order, so waiting
// for the outstanding ones (lgkmcnt) makes every earlier LDS write visible to
// every lane's later reads; the memory clobber keeps the compiler from moving
// accesses across. Global memory operations in flight are not waited for.
__device__ inline void wave_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Cell i's candidates staged in LDS for the join kernels (one wave per block,
// the whole wave calls; m <= 64, so nc <= 128 and lane l holds candidates l and
// l + 64). Loads that do not depend on each other are issued together, so a cell
// waits on two memory round trips, not one per 64 items: (1) its counts and both
// lists, read whole (m entries each; entries past the counts are read and
// dropped), (2) the candidates' thresholds and rows, STAGE_BATCH 16-byte loads
// per lane in flight. Candidate s (new list, then old list) goes to s_id[s],
// s_thr[s] and row s of `rows` (stride RS floats); with `xi`, row i goes there.
// Returns false, with nothing staged, when the cell has no new candidate or its
// nc candidates exceed the rcap rows the caller holds (another launch takes it).
// The caller synchronises before reading what was staged.
constexpr int STAGE_BATCH = 8;
template <int D>
__device__ inline bool stage_cell(const float* __restrict__ Xp, uint32_t d, uint32_t dp, uint32_t RS,
                                  const uint32_t* __restrict__ newl, const uint32_t* __restrict__ oldl,
                                  const uint32_t* __restrict__ ncnt, uint32_t m,
                                  const float* __restrict__ thr, uint32_t i, uint32_t rcap,
                                  uint32_t* s_id, float* s_thr, float* rows, float* xi, uint32_t& nn,
                                  uint32_t& no) {
  const uint32_t l = (uint32_t)lane_id();
  const size_t lb = (size_t)i * m;
  // loads are unconditional, on clamped in-bounds addresses, so the compiler
  // issues them back to back and waits once
  const uint32_t lm = l < m ? l : m - 1;
  const uint32_t nv = newl[lb + lm];
  const uint32_t ov = oldl[lb + lm];
  const uint32_t cnt = ncnt[i];
  nn = cnt & 0xFFFF;
  no = cnt >> 16;
  const uint32_t nc = nn + no;
  if (nn == 0 || nc > rcap) return false;
  // every lane takes part in each shuffle; the selects come after
  const uint32_t o0 = (uint32_t)__shfl((int)ov, (int)((l - nn) & (WAVE - 1)), WAVE);
  const uint32_t v0 = l < nn ? nv : o0;
  const uint32_t v1 = (uint32_t)__shfl((int)ov, (int)((l + WAVE - nn) & (WAVE - 1)), WAVE);
  const float t0 = thr[l < nc ? v0 : i];
  const float t1 = thr[l + WAVE < nc ? v1 : i];
  // Rows are copied in passes of WAVE / lpr rows, lpr lanes per row: lane l
  // moves 16-byte chunk c = l % lpr of row s = pass * (WAVE / lpr) + l / lpr
  // (chunks past c4 idle), so no lane divides by the row length. A pass never
  // straddles row 64, so its rows' ids come from one of v0 / v1 (uniform).
  const uint32_t c4 = D > 0 ? (uint32_t)(D + 3) / 4 : (d + 3) / 4;
  const uint32_t lpr = c4 <= 16 ? 16u : (uint32_t)WAVE;
  const uint32_t rpp = (uint32_t)WAVE / lpr;
  const uint32_t lr = l / lpr, c = l % lpr;
  const bool cact = c < c4;
  const uint32_t cc = cact ? c : 0u;
  const uint32_t nrows = nc + (xi != nullptr ? 1u : 0u);
  const uint32_t npass = (nrows + rpp - 1) / rpp;
  // LDS destinations as 32-bit float offsets from `rows` (x_i's row at xoff)
  const uint32_t xoff = xi != nullptr ? (uint32_t)(xi - rows) : 0u;
  for (uint32_t p0 = 0; p0 < npass; p0 += STAGE_BATCH) {
    // passes past npass (uniform) are skipped
    float4 v[STAGE_BATCH];
#pragma unroll
    for (int u = 0; u < STAGE_BATCH; ++u) {
      if (p0 + (uint32_t)u >= npass) break;
      const uint32_t base = (p0 + (uint32_t)u) * rpp;  // uniform
      const uint32_t s = base + lr;
      // every lane takes part in the permute (byte address of lane s mod 64); the
      // select comes after
      const uint32_t a = (uint32_t)__builtin_amdgcn_ds_bpermute(
          (int)(((base & (WAVE - 1)) + lr) << 2), (int)(base < (uint32_t)WAVE ? v0 : v1));
      const uint32_t src = s < nc ? a : i;
      v[u] = *reinterpret_cast<const float4*>(Xp + (size_t)src * dp + 4 * cc);
    }
    // the values are pinned here, so no load is sunk into its (conditional) store
#pragma unroll
    for (int u = 0; u < STAGE_BATCH; ++u) {
      if (p0 + (uint32_t)u >= npass) break;
      asm volatile("" : "+v"(v[u].x), "+v"(v[u].y), "+v"(v[u].z), "+v"(v[u].w));
    }
#pragma unroll
    for (int u = 0; u < STAGE_BATCH; ++u) {
      if (p0 + (uint32_t)u >= npass) break;
      const uint32_t s = (p0 + (uint32_t)u) * rpp + lr;
      if (cact && s < nrows)
        *reinterpret_cast<float4*>(rows + ((s < nc ? s * RS : xoff) + 4 * c)) = v[u];
    }
  }
  if (l < nc) {
    s_id[l] = v0;
    s_thr[l] = t0;
  }
  if (l + WAVE < nc) {
    s_id[l + WAVE] = v1;
    s_thr[l + WAVE] = t1;
  }
  return true;
}

// A join launch holds rcap candidate rows per block in LDS. Where 2m > rcap,
// the launch over every cell skips the cells with more candidates, and a second
// launch, with room for 2m rows, takes those from a list in cell order, one per
// block. Both map blocks to cells with xcd_row(), so each XCD joins a
// contiguous run of cells (and its L2 holds their neighbourhood). Cells are
// independent (every proposal goes through the order-free bucket insert).
struct JoinCells {
  const uint32_t* list;  // NULL: cells 0..n-1
  uint32_t count;        // cells in the list
};
// This block's cell, or false when it has none (block-uniform).
__device__ inline bool join_cell_of_block(const JoinCells& jc, uint32_t n, uint32_t& i) {
  const uint32_t e = xcd_row();
  if (jc.list) {
    if (e >= jc.count) return false;
    i = jc.list[e];
    return true;
  }
  i = e;
  return i < n;
}

// The join kernels run one wave per block, so __syncthreads() is a wave barrier
// and every call below is made by the whole wave. Capacity: < 64 entries are
// held between calls and one append adds at most 64 (callers flush after each
// append).
constexpr uint32_t PB_CAP = 128;
__device__ inline void pb_append(uint32_t* pb_row, uint64_t* pb_key, uint32_t& cnt, bool has,
                                 uint32_t row, uint64_t key, uint64_t& sent) {
  const uint64_t mask = __ballot(has);
  if (has) {
    const uint32_t pos = cnt + (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
    pb_row[pos] = row;
    pb_key[pos] = key;
  }
  cnt += (uint32_t)__popcll(mask);
  sent += (uint64_t)__popcll(mask);
}
// Sends every full batch of 64 (all entries when `all`), keeping the rest.
__device__ inline void pb_flush(uint32_t* pb_row, uint64_t* pb_key, uint32_t& cnt, bool all,
                                uint64_t* __restrict__ slots, uint8_t* __restrict__ touched,
                                uint32_t S, const SlotHash& sh) {
  if (cnt < (all ? 1u : (uint32_t)WAVE)) return;
  const uint32_t l = (uint32_t)lane_id();
  __syncthreads();
  uint32_t done = 0;
  for (; done + WAVE <= cnt; done += WAVE)
    propose(slots, touched, pb_row[done + l], pb_key[done + l], S, sh);
  const uint32_t rest = cnt - done;
  if (all) {
    if (l < rest) propose(slots, touched, pb_row[done + l], pb_key[done + l], S, sh);
    cnt = 0;
    return;
  }
  uint32_t rr = 0;
  uint64_t rk = 0;
  if (l < rest) {
    rr = pb_row[done + l];
    rk = pb_key[done + l];
  }
  __syncthreads();
  if (l < rest) {
    pb_row[l] = rr;
    pb_key[l] = rk;
  }
  __syncthreads();
  cnt = rest;
}

}  // namespace psq
