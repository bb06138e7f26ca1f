// Fuzzy connectivities of a kNN distance graph on the GPU
// (crates/polariseq-core/src/umap.rs:130-139): membership strengths per row
// (umap.rs:551-608), the transpose by counting sort (umap.rs:357-381) and the
// per-row fuzzy union of a row's own entries with the transposed ones
// (umap.rs:488-537).
//
// Row i of the symmetric graph is built from a combined segment of
// comb[indptr[i] + t_ptr[i], indptr[i+1] + t_ptr[i+1]): the row's own entries,
// then the entries of rows that name i (t_ptr = running sum of in-degrees).
// Each entry is a sort key (column << 32) | (tag << 31) | weight bits, tag 0
// for own, 1 for transposed, so ascending keys give column order with, per
// column, own entries before transposed ones: the order merge_fuzzy_row folds
// them in. Weights are >= +0 or NaN, so the sign bit carries no information
// (a NaN weight folds to NaN and is dropped either way).
#include "psq_kernels.hpp"
#include "psq_partition.hpp"

namespace psq {

namespace {

__device__ inline uint64_t fz_key(uint32_t col, uint32_t tag, float w) {
  return ((uint64_t)col << 32) | ((uint64_t)tag << 31) | (__float_as_uint(w) & 0x7fffffffu);
}
__device__ inline uint32_t fz_col(uint64_t k) { return (uint32_t)(k >> 32); }
__device__ inline float fz_w(uint64_t k) { return __uint_as_float((uint32_t)k & 0x7fffffffu); }

__device__ inline double fz_term(double d, double rho, double mid) {
  const double x = d - rho;
  return x > 0.0 ? exp(-(x / mid)) : 1.0;
}

// umap.rs:551-608 for one row of `len` distances read through get(j).
// The sums run left to right in f64, as Rust's iterator sum does.
template <typename Get, typename Put>
__device__ inline void fz_smooth(Get get, uint32_t len, Put put) {
  const double target = log2((double)(len + 1));
  double rho = 0.0;  // local_connectivity = 1: the nearest non-zero distance
  for (uint32_t j = 0; j < len; ++j) {
    const double v = get(j);
    if (v > 0.0) {
      rho = v;
      break;
    }
  }
  double lo = 0.0, hi = INFINITY, mid = 1.0;
  for (int it = 0; it < 64; ++it) {
    double psum = 0.0;
    for (uint32_t j = 0; j < len; ++j) psum += fz_term(get(j), rho, mid);
    if (fabs(psum - target) < 1e-5) break;
    if (psum > target) {
      hi = mid;
      mid = (lo + hi) / 2.0;
    } else {
      lo = mid;
      mid = isinf(hi) ? mid * 2.0 : (lo + hi) / 2.0;
    }
  }
  double sum = 0.0;
  for (uint32_t j = 0; j < len; ++j) sum += get(j);
  const double mean = sum / (double)(len + 1);  // the self distance of zero counts
  const double sigma = rho > 0.0 ? fmax(mid, 1e-3 * mean) : mid;
  for (uint32_t j = 0; j < len; ++j) {
    const double x = get(j) - rho;
    put(j, (x <= 0.0 || sigma == 0.0) ? 1.0f : (float)exp(-(x / sigma)));
  }
}

// Fold the sorted entries s[0, len) column by column (umap.rs:517-535) and
// compact the kept ones, as (column << 32) | weight bits, into out[0, total).
// blockDim.x == FZ_BLOCK; every thread of the block calls it. s may be in LDS
// or global memory.
__device__ inline uint32_t fz_fold_block(const uint64_t* s, uint32_t len, uint64_t* __restrict__ out,
                                         uint32_t* s_w) {
  const int tid = threadIdx.x, l = tid & (WAVE - 1), w = tid / WAVE;
  constexpr int NW = FZ_BLOCK / WAVE;
  uint32_t cnt = 0;
  for (uint32_t b = 0; b < len; b += FZ_BLOCK) {
    const uint32_t e = b + tid;
    bool keep = false;
    uint64_t outk = 0;
    if (e < len) {
      const uint32_t col = fz_col(s[e]);
      if (e == 0 || fz_col(s[e - 1]) != col) {
        float acc = 0.0f;
        for (uint32_t t = e; t < len && fz_col(s[t]) == col; ++t) {
          const float x = fz_w(s[t]);
          acc = acc + x - acc * x;
        }
        keep = acc > 1e-5f;
        outk = ((uint64_t)col << 32) | __float_as_uint(acc);
      }
    }
    const uint64_t mask = __ballot(keep);
    if (l == 0) s_w[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const uint32_t c = s_w[q];
      before += q < w ? c : 0u;
      all += c;
    }
    if (keep) out[cnt + before + (uint32_t)__popcll(mask & ((1ull << l) - 1))] = outk;
    cnt += all;
    __syncthreads();
  }
  return cnt;
}


// One bitonic compare-exchange step over pairs t of a P-long sequence.
__device__ inline void fz_bitonic_pair(uint64_t* s, uint32_t t, uint32_t size, uint32_t stride) {
  const uint32_t lo = 2 * t - (t & (stride - 1));
  const uint32_t hi = lo + stride;
  const bool up = (lo & size) == 0;
  const uint64_t a = s[lo], b = s[hi];
  if ((a > b) == up) {
    s[lo] = b;
    s[hi] = a;
  }
}

}  // namespace

// Row pointers non-decreasing and every index < n; bit 1 / bit 2 of *bad otherwise.
__global__ void k_fuzzy_check(const uint32_t* __restrict__ indptr, uint32_t n,
                              const uint32_t* __restrict__ indices, uint64_t m,
                              uint32_t* __restrict__ bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t flags = 0;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m || t < n; t += stride) {
    if (t < n && indptr[t + 1] < indptr[t]) flags |= 1u;
    if (t < m && indices[t] >= n) flags |= 2u;
  }
  if (flags) atomicOr(bad, flags);
}

// umap.rs:343-354: one thread per row, distances held in registers when the
// row has at most FZ_KR entries. Each strength w of edge (i -> j) is written as
// the transposed entry it becomes, fz_key(i, 1, w), at the edge's position.
__global__ __launch_bounds__(256) void k_fuzzy_strengths(const uint32_t* __restrict__ indptr,
                                                         const float* __restrict__ dists, uint32_t n,
                                                         uint64_t* __restrict__ pay0) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = indptr[i];
  const uint32_t len = indptr[i + 1] - s;
  if (len == 0) return;
  const float* dr = dists + s;
  uint64_t* pr = pay0 + s;
  if (len <= FZ_KR) {
    double d[FZ_KR];
#pragma unroll
    for (int j = 0; j < FZ_KR; ++j) d[j] = (uint32_t)j < len ? (double)dr[j] : 0.0;
    // The loops of fz_smooth over j < len, written over the unrolled range.
    const double target = log2((double)(len + 1));
    double rho = 0.0;
    bool found = false;
#pragma unroll
    for (int j = 0; j < FZ_KR; ++j)
      if (!found && (uint32_t)j < len && d[j] > 0.0) {
        rho = d[j];
        found = true;
      }
    double lo = 0.0, hi = INFINITY, mid = 1.0;
    for (int it = 0; it < 64; ++it) {
      double psum = 0.0;
#pragma unroll
      for (int j = 0; j < FZ_KR; ++j)
        if ((uint32_t)j < len) psum += fz_term(d[j], rho, mid);
      if (fabs(psum - target) < 1e-5) break;
      if (psum > target) {
        hi = mid;
        mid = (lo + hi) / 2.0;
      } else {
        lo = mid;
        mid = isinf(hi) ? mid * 2.0 : (lo + hi) / 2.0;
      }
    }
    double sum = 0.0;
#pragma unroll
    for (int j = 0; j < FZ_KR; ++j)
      if ((uint32_t)j < len) sum += d[j];
    const double mean = sum / (double)(len + 1);
    const double sigma = rho > 0.0 ? fmax(mid, 1e-3 * mean) : mid;
#pragma unroll
    for (int j = 0; j < FZ_KR; ++j)
      if ((uint32_t)j < len) {
        const double x = d[j] - rho;
        pr[j] = fz_key(i, 1, (x <= 0.0 || sigma == 0.0) ? 1.0f : (float)exp(-(x / sigma)));
      }
  } else {
    fz_smooth([&](uint32_t j) { return (double)dr[j]; }, len,
              [&](uint32_t j, float v) { pr[j] = fz_key(i, 1, v); });
  }
}

// ── transpose by bucket (umap.rs:357-381) ──
// Every edge (i -> j, w) becomes the transposed entry (column i, tag 1, w) of
// row j through the two staged partitions of psq_partition.hpp: level 1 here
// (edges -> super-buckets), level 2 by k_part_l2_* (-> buckets of
// FZ_BUCKET_ROWS rows). k_fz_bucket_merge then counting-sorts one bucket's own
// and transposed entries by row in LDS and merges every row of the bucket.

// Level-1 count: edges per target super-bucket (at most SBM <= 1024 of them).
template <int SBM>
__global__ __launch_bounds__(1024) void k_fz_l1_count(const uint32_t* __restrict__ indices,
                                                      uint64_t m, uint32_t sshift, uint32_t ns,
                                                      uint32_t* __restrict__ stot) {
  __shared__ uint32_t h[SBM];
  const uint32_t tid = threadIdx.x;
  if (tid < ns) h[tid] = 0;
  __syncthreads();
  const uint64_t e0 = (uint64_t)blockIdx.x * PART_C1_CHUNK;
  const uint64_t e1 = min(m, e0 + PART_C1_CHUNK);
  for (uint64_t e = e0 + tid; e < e1; e += blockDim.x) atomicAdd(&h[indices[e] >> sshift], 1u);
  __syncthreads();
  if (tid < ns && h[tid]) atomicAdd(&stot[tid], h[tid]);
}

// Level-1 placement: pay1 = the transposed entry key, pj1 = the full target
// row j, in super-bucket order.
template <int SBM>
__global__ __launch_bounds__(1024) void k_fz_l1_scatter(
    const uint64_t* __restrict__ pay0, const uint32_t* __restrict__ indices, uint64_t m,
    uint32_t sshift, uint32_t ns, const uint32_t* __restrict__ sstart, uint32_t* __restrict__ sfill,
    uint64_t* __restrict__ pay1, uint32_t* __restrict__ pj1) {
  __shared__ uint64_t s_key[PART_P1_CHUNK];
  __shared__ uint32_t s_tgt[PART_P1_CHUNK];
  __shared__ uint32_t s_cnt[SBM + 1], s_base[SBM], s_part[1024 / WAVE];
  const uint64_t e0 = (uint64_t)blockIdx.x * PART_P1_CHUNK;
  const uint32_t cnt = (uint32_t)min((uint64_t)PART_P1_CHUNK, m - e0);
  const uint64_t* kin = pay0 + e0;
  const uint32_t* tin = indices + e0;
  stage_partition<PART_P1_CHUNK / 1024>(
      [&](uint32_t e, uint64_t& kk, uint32_t& tt) {
        kk = kin[e];
        tt = tin[e];
        return true;
      },
      cnt, ns, 0u, [&](uint32_t j) { return j >> sshift; }, sstart, sfill, pay1, pj1,
      [](uint32_t j) { return j; }, s_key, s_tgt, s_cnt, s_base, s_part);
}
template __global__ void k_fz_l1_count<PART_SB_MAX>(const uint32_t*, uint64_t, uint32_t, uint32_t, uint32_t*);
template __global__ void k_fz_l1_count<PART_SB_WIDE>(const uint32_t*, uint64_t, uint32_t, uint32_t, uint32_t*);
template __global__ void k_fz_l1_scatter<PART_SB_MAX>(const uint64_t*, const uint32_t*, uint64_t, uint32_t,
                                                      uint32_t, const uint32_t*, uint32_t*, uint64_t*,
                                                      uint32_t*);
template __global__ void k_fz_l1_scatter<PART_SB_WIDE>(const uint64_t*, const uint32_t*, uint64_t, uint32_t,
                                                       uint32_t, const uint32_t*, uint32_t*, uint64_t*,
                                                       uint32_t*);

// One block per bucket of FZ_BUCKET_ROWS rows: the bucket's own entries (a
// contiguous CSR range) and its transposed entries are counting-sorted by row
// into one array, in LDS when they fit (FZ_BCAP), else in the bucket's own
// region of comb. t_ptr of the bucket's rows is written. Then one wave per row
// sorts the row's segment in registers, writes it back, folds and compacts it
// to the start of the row's combined segment in comb; olen = kept entries.
// Rows of more than 256 combined entries are written unsorted to their comb
// segment and listed in big[] for k_fuzzy_merge_mid.
__global__ __launch_bounds__(FZ_MB, 2) void k_fz_bucket_merge(
    const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
    const uint64_t* __restrict__ pay0, uint32_t n, uint32_t nb, const uint32_t* __restrict__ bstart,
    const uint64_t* __restrict__ pay2, const uint16_t* __restrict__ pl2,
    uint32_t* __restrict__ t_ptr, uint64_t* __restrict__ comb, uint32_t* __restrict__ olen,
    uint32_t* __restrict__ big, uint32_t* __restrict__ nbig) {
  extern __shared__ uint64_t lds[];
  uint64_t* s_srt = lds;  // FZ_BCAP
  __shared__ uint32_t s_ip[FZ_BUCKET_ROWS + 1], s_off[FZ_BUCKET_ROWS + 1], s_cur[FZ_BUCKET_ROWS];
  __shared__ uint32_t s_part[FZ_MB / WAVE];
  __shared__ uint32_t s_short[FZ_BUCKET_ROWS], s_long[FZ_BUCKET_ROWS], s_ncls[2];
  const uint32_t tid = threadIdx.x, l = tid & (WAVE - 1), wv = tid / WAVE;
  const uint32_t b = blockIdx.x;
  if (tid < 2) s_ncls[tid] = 0;
  const uint32_t row0 = b * FZ_BUCKET_ROWS, rows = min((uint32_t)FZ_BUCKET_ROWS, n - row0);
  const uint32_t p0 = bstart[b], p1 = bstart[b + 1], c = p1 - p0;
  for (uint32_t r = tid; r <= rows; r += FZ_MB) s_ip[r] = indptr[row0 + r];
  __syncthreads();
  const uint32_t a0 = s_ip[0], own_all = s_ip[rows] - a0, total = own_all + c;
  const uint64_t cbase = (uint64_t)a0 + p0;  // the bucket's region of comb
  for (uint32_t r = tid; r < rows; r += FZ_MB) s_off[r] = s_ip[r + 1] - s_ip[r];
  __syncthreads();
  for (uint32_t t = tid; t < c; t += FZ_MB) atomicAdd(&s_off[pl2[p0 + t]], 1u);
  __syncthreads();
  block_scan4(s_off, rows, s_part);
  for (uint32_t r = tid; r < rows; r += FZ_MB) {
    const uint32_t own_r = s_ip[r + 1] - s_ip[r];
    s_cur[r] = s_off[r] + own_r;
    t_ptr[row0 + r] = p0 + s_off[r] - (s_ip[r] - a0);
  }
  if (b == nb - 1 && tid == 0) t_ptr[n] = p1;
  // The rest runs on the sorted array in LDS or, for a bucket too large for
  // it, in the bucket's comb region; each call below sees one address space.
  auto rest = [&](uint64_t* srt) __attribute__((always_inline)) {
    // own entries: the bucket's contiguous CSR range, each entry's row found by
    // binary search over the row pointers
    for (uint32_t t = tid; t < own_all; t += FZ_MB) {
      const uint32_t q = a0 + t;
      uint32_t lo = 0, hi = rows;  // s_ip[lo] <= q < s_ip[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_ip[mid] <= q) lo = mid;
        else hi = mid;
      }
      srt[s_off[lo] + (q - s_ip[lo])] = ((uint64_t)indices[q] << 32) | ((uint32_t)pay0[q] & 0x7fffffffu);
    }
    __syncthreads();
    for (uint32_t t = tid; t < c; t += FZ_MB) srt[atomicAdd(&s_cur[pl2[p0 + t]], 1u)] = pay2[p0 + t];
    __syncthreads();
    // Rows of at most 32 entries are taken in pairs, a half wave each; longer
    // rows one per wave.
    constexpr uint32_t NW = FZ_MB / WAVE;
    for (uint32_t r = tid; r < rows; r += FZ_MB) {
      const uint32_t len = s_off[r + 1] - s_off[r];
      if (len <= WAVE / 2) s_short[atomicAdd(&s_ncls[0], 1u)] = r;
      else s_long[atomicAdd(&s_ncls[1], 1u)] = r;
    }
    __syncthreads();
    const uint32_t nshort = s_ncls[0], nlong = s_ncls[1];
    for (uint32_t pr = wv; 2 * pr < nshort; pr += NW) {
      const uint32_t h = l / (WAVE / 2), el = l % (WAVE / 2);
      const bool has = 2 * pr + h < nshort;
      const uint32_t r = has ? s_short[2 * pr + h] : 0u;
      const uint32_t so = has ? s_off[r] : 0u;
      const uint32_t len = has ? s_off[r + 1] - so : 0u;
      uint64_t* seg = srt + so;
      const uint64_t v = wave_sort32x2(el < len ? seg[el] : ~0ull);
      if (el < len) seg[el] = v;
      wave_sync();
      bool keep = false;
      uint64_t outk = 0;
      if (el < len) {
        const uint32_t col = fz_col(seg[el]);
        if (el == 0 || fz_col(seg[el - 1]) != col) {
          float acc = 0.0f;
          for (uint32_t t = el; t < len && fz_col(seg[t]) == col; ++t) {
            const float x = fz_w(seg[t]);
            acc = acc + x - acc * x;
          }
          keep = acc > 1e-5f;
          outk = ((uint64_t)col << 32) | __float_as_uint(acc);
        }
      }
      const uint32_t hm = (uint32_t)(__ballot(keep) >> (h * (WAVE / 2)));
      wave_sync();
      if (keep) comb[cbase + so + (uint32_t)__popc(hm & ((1u << el) - 1u))] = outk;
      if (el == 0 && has) olen[row0 + r] = (uint32_t)__popc(hm);
      wave_sync();
    }
    for (uint32_t li = wv; li < nlong; li += NW) {
      const uint32_t r = s_long[li];
      const uint32_t row = row0 + r;
      const uint32_t so = s_off[r], len = s_off[r + 1] - so;
      uint64_t* seg = srt + so;
      const uint64_t coff = cbase + so;
      if (len > 4 * WAVE) {
        if (srt != comb + cbase)
          for (uint32_t q = l; q < len; q += WAVE) comb[coff + q] = seg[q];
        if (l == 0) big[atomicAdd(nbig, 1u)] = row;
        continue;
      }
      if (len <= WAVE) {
        const uint64_t v = wave_sort64((uint32_t)l < len ? seg[l] : ~0ull);
        if ((uint32_t)l < len) seg[l] = v;
      } else {
        uint64_t v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t e = q * WAVE + l;
          v[q] = e < len ? seg[e] : ~0ull;
        }
        wave_sort_n<4>(v);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if ((uint32_t)(q * WAVE + l) < len) seg[q * WAVE + l] = v[q];
      }
      wave_sync();
      uint32_t kept = 0;
      for (uint32_t base = 0; base < len; base += WAVE) {
        const uint32_t e = base + l;
        bool keep = false;
        uint64_t outk = 0;
        if (e < len) {
          const uint32_t col = fz_col(seg[e]);
          if (e == 0 || fz_col(seg[e - 1]) != col) {
            float acc = 0.0f;
            for (uint32_t t = e; t < len && fz_col(seg[t]) == col; ++t) {
              const float x = fz_w(seg[t]);
              acc = acc + x - acc * x;
            }
            keep = acc > 1e-5f;
            outk = ((uint64_t)col << 32) | __float_as_uint(acc);
          }
        }
        const uint64_t mask = __ballot(keep);
        wave_sync();
        if (keep) comb[coff + kept + (uint32_t)__popcll(mask & ((1ull << l) - 1))] = outk;
        kept += (uint32_t)__popcll(mask);
        wave_sync();
      }
      if (l == 0) olen[row] = kept;
    }
  };
  if (total <= (uint32_t)FZ_BCAP) rest(s_srt);
  else rest(comb + cbase);
}

// Rows of 257 .. FZ_MID_CAP combined entries: one block per row (looping over
// the list), bitonic sort in LDS, fold and compact. Longer rows go to giant[].
__global__ __launch_bounds__(FZ_BLOCK) void k_fuzzy_merge_mid(
    const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ t_ptr,
    uint64_t* __restrict__ comb, uint32_t* __restrict__ olen, const uint32_t* __restrict__ big,
    const uint32_t* __restrict__ nbig, uint32_t* __restrict__ giant, uint32_t* __restrict__ ngiant) {
  extern __shared__ uint64_t s[];  // FZ_MID_CAP entries
  __shared__ uint32_t s_w[FZ_BLOCK / WAVE];
  const uint32_t nb = *nbig;
  const int tid = threadIdx.x;
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const uint32_t i = big[b];
    const uint64_t off = (uint64_t)indptr[i] + t_ptr[i];
    const uint32_t len = (indptr[i + 1] - indptr[i]) + (t_ptr[i + 1] - t_ptr[i]);
    if (len > FZ_MID_CAP) {
      if (tid == 0) giant[atomicAdd(ngiant, 1u)] = i;
      continue;
    }
    uint32_t P = 1;
    while (P < len) P <<= 1;
    for (uint32_t t = tid; t < P; t += FZ_BLOCK) s[t] = t < len ? comb[off + t] : ~0ull;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1)
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t t = tid; t < P / 2; t += FZ_BLOCK) fz_bitonic_pair(s, t, size, stride);
        __syncthreads();
      }
    const uint32_t c = fz_fold_block(s, len, comb + off, s_w);
    if (tid == 0) olen[i] = c;
    __syncthreads();
  }
}

// Giant rows: lengths, padded copies, global bitonic steps, fold.
__global__ void k_fuzzy_row_len(const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ t_ptr,
                                const uint32_t* __restrict__ rows, uint32_t nrows,
                                uint32_t* __restrict__ len) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const uint32_t i = rows[r];
  len[r] = (indptr[i + 1] - indptr[i]) + (t_ptr[i + 1] - t_ptr[i]);
}

__global__ void k_fuzzy_gather_pad(const uint32_t* __restrict__ indptr,
                                   const uint32_t* __restrict__ t_ptr,
                                   const uint32_t* __restrict__ rows, uint32_t nrows, uint32_t P,
                                   const uint64_t* __restrict__ comb, uint64_t* __restrict__ tmp) {
  const uint64_t total = (uint64_t)nrows * P;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const uint32_t r = (uint32_t)(t / P), e = (uint32_t)(t % P);
    const uint32_t i = rows[r];
    const uint32_t len = (indptr[i + 1] - indptr[i]) + (t_ptr[i + 1] - t_ptr[i]);
    tmp[t] = e < len ? comb[(uint64_t)indptr[i] + t_ptr[i] + e] : ~0ull;
  }
}

__global__ void k_fuzzy_bitonic_step(uint64_t* __restrict__ tmp, uint32_t nrows, uint32_t P,
                                     uint32_t size, uint32_t stride) {
  const uint64_t half = P / 2, total = (uint64_t)nrows * half;
  const uint64_t gs = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gs) {
    const uint64_t r = t / half;
    fz_bitonic_pair(tmp + r * P, (uint32_t)(t % half), size, stride);
  }
}

__global__ __launch_bounds__(FZ_BLOCK) void k_fuzzy_fold_rows(const uint32_t* __restrict__ indptr,
                                                              const uint32_t* __restrict__ t_ptr,
                                                              const uint32_t* __restrict__ rows,
                                                              uint32_t nrows, uint32_t P,
                                                              const uint64_t* __restrict__ tmp,
                                                              uint64_t* __restrict__ comb,
                                                              uint32_t* __restrict__ olen) {
  __shared__ uint32_t s_w[FZ_BLOCK / WAVE];
  for (uint32_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    const uint32_t i = rows[r];
    const uint64_t off = (uint64_t)indptr[i] + t_ptr[i];
    const uint32_t len = (indptr[i + 1] - indptr[i]) + (t_ptr[i + 1] - t_ptr[i]);
    const uint32_t c = fz_fold_block(tmp + (uint64_t)r * P, len, comb + off, s_w);
    if (threadIdx.x == 0) olen[i] = c;
  }
}

// Sum of the kept counts in 64 bits (the u32 running sum wraps past 2^32).
__global__ void k_fuzzy_total(const uint32_t* __restrict__ olen, uint32_t n,
                              unsigned long long* __restrict__ total) {
  unsigned long long v = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) v += olen[t];
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += shfl_xor64(v, o);
  if ((threadIdx.x & (WAVE - 1)) == 0 && v) atomicAdd(total, v);
}

// FZ_EMIT_LANES lanes per row copy the compacted entries into the packed output.
__global__ __launch_bounds__(256) void k_fuzzy_emit(const uint32_t* __restrict__ indptr,
                                                    const uint32_t* __restrict__ t_ptr, uint32_t n,
                                                    const uint64_t* __restrict__ comb,
                                                    const uint32_t* __restrict__ optr,
                                                    uint32_t* __restrict__ out_ix,
                                                    float* __restrict__ out_w) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = g / FZ_EMIT_LANES, l = g % FZ_EMIT_LANES;
  if (i >= n) return;
  const uint64_t off = (uint64_t)indptr[i] + t_ptr[i];
  const uint32_t o = optr[i], c = optr[i + 1] - o;
  for (uint32_t t = l; t < c; t += FZ_EMIT_LANES) {
    const uint64_t k = comb[off + t];
    out_ix[o + t] = (uint32_t)(k >> 32);
    out_w[o + t] = __uint_as_float((uint32_t)k);
  }
}

}  // namespace psq
