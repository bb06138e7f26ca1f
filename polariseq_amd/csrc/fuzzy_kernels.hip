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
template <typename Get>
__device__ inline void fz_smooth(Get get, uint32_t len, float* __restrict__ wr) {
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
    wr[j] = (x <= 0.0 || sigma == 0.0) ? 1.0f : (float)exp(-(x / sigma));
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
// row has at most FZ_KR entries.
__global__ __launch_bounds__(256) void k_fuzzy_strengths(const uint32_t* __restrict__ indptr,
                                                         const float* __restrict__ dists, uint32_t n,
                                                         float* __restrict__ w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = indptr[i];
  const uint32_t len = indptr[i + 1] - s;
  if (len == 0) return;
  const float* dr = dists + s;
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
    float* wr = w + s;
#pragma unroll
    for (int j = 0; j < FZ_KR; ++j)
      if ((uint32_t)j < len) {
        const double x = d[j] - rho;
        wr[j] = (x <= 0.0 || sigma == 0.0) ? 1.0f : (float)exp(-(x / sigma));
      }
  } else {
    fz_smooth([&](uint32_t j) { return (double)dr[j]; }, len, w + s);
  }
}

// ── transpose by bucket (umap.rs:357-381) ──
// Rows are grouped in buckets of 2^shift consecutive rows. k_fz_bucket_count
// counts the edges that point into each bucket, k_fz_bucket_scatter writes each
// edge (i -> j, w) into its target bucket's region as the transposed entry of
// row j, and k_fz_bucket_merge sorts one bucket's entries by row in LDS and
// merges every row of the bucket. Global atomics are one per (tile, bucket);
// the per-edge counting runs on LDS atomics.

// Per tile of FZ_TILE_ROWS rows: edges per target bucket, added to btot.
__global__ __launch_bounds__(1024) void k_fz_bucket_count(const uint32_t* __restrict__ indptr,
                                                          const uint32_t* __restrict__ indices,
                                                          uint32_t n, uint32_t shift, uint32_t nb,
                                                          uint32_t* __restrict__ btot) {
  __shared__ uint32_t h[FZ_NB_MAX];
  const uint32_t tid = threadIdx.x;
  for (uint32_t b = tid; b < nb; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const uint32_t r0 = blockIdx.x * FZ_TILE_ROWS, r1 = min(n, r0 + FZ_TILE_ROWS);
  const uint32_t e0 = indptr[r0], e1 = indptr[r1];
  for (uint32_t q = e0 + tid; q < e1; q += blockDim.x) atomicAdd(&h[indices[q] >> shift], 1u);
  __syncthreads();
  for (uint32_t b = tid; b < nb; b += blockDim.x)
    if (h[b]) atomicAdd(&btot[b], h[b]);
}

// Same tiles: reserve each bucket's share with one global atomic, then place
// the tile's edges: pay = transposed entry key of row j (column i, tag 1),
// pj = j's row within its bucket.
__global__ __launch_bounds__(1024) void k_fz_bucket_scatter(
    const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
    const float* __restrict__ w, uint32_t n, uint32_t shift, uint32_t nb,
    const uint32_t* __restrict__ bstart, uint32_t* __restrict__ bfill, uint64_t* __restrict__ pay,
    uint32_t* __restrict__ pj) {
  __shared__ uint32_t h[FZ_NB_MAX];
  const uint32_t tid = threadIdx.x;
  for (uint32_t b = tid; b < nb; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const uint32_t r0 = blockIdx.x * FZ_TILE_ROWS, r1 = min(n, r0 + FZ_TILE_ROWS);
  const uint32_t e0 = indptr[r0], e1 = indptr[r1];
  for (uint32_t q = e0 + tid; q < e1; q += blockDim.x) atomicAdd(&h[indices[q] >> shift], 1u);
  __syncthreads();
  for (uint32_t b = tid; b < nb; b += blockDim.x)
    if (h[b]) h[b] = bstart[b] + atomicAdd(&bfill[b], h[b]);
  __syncthreads();
  const uint32_t l = tid % FZ_SCATTER_LANES, per = blockDim.x / FZ_SCATTER_LANES;
  const uint32_t mask = (1u << shift) - 1u;
  for (uint32_t r = r0 + tid / FZ_SCATTER_LANES; r < r1; r += per) {
    const uint32_t s = indptr[r], e = indptr[r + 1];
    for (uint32_t q = s + l; q < e; q += FZ_SCATTER_LANES) {
      const uint32_t j = indices[q];
      const uint32_t pos = atomicAdd(&h[j >> shift], 1u);
      pay[pos] = fz_key(r, 1, w[q]);
      pj[pos] = j & mask;
    }
  }
}

namespace {

// Exclusive scan in place of v[0, L) by the whole block (blockDim.x == FZ_MB);
// v[L] = total. v may be in LDS or global memory.
__device__ inline void fz_block_scan(uint32_t* v, uint32_t L, uint32_t* s_part) {
  const uint32_t tid = threadIdx.x, l = tid & (WAVE - 1), wv = tid / WAVE;
  constexpr uint32_t PER = 4, NWV = FZ_MB / WAVE;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < L; base += FZ_MB * PER) {
    const uint32_t i0 = base + tid * PER;
    uint32_t x[PER], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
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
      uint32_t t = l < NWV ? s_part[l] : 0u;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(t, o, WAVE);
        if (l >= (uint32_t)o) t += y;
      }
      if (l < NWV) s_part[l] = t;
    }
    __syncthreads();
    uint32_t run = carry + (wv > 0 ? s_part[wv - 1] : 0u) + (inc - sum);
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k)
      if (i0 + k < L) {
        v[i0 + k] = run;
        run += x[k];
      }
    carry += s_part[NWV - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) v[L] = carry;
  __syncthreads();
}

__device__ inline void fz_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace

// One block per bucket: counting sort of the bucket's transposed entries by row
// (in LDS when they fit, else in srt_g at the bucket's region), t_ptr of its
// rows, then one wave per row: the row's own entries and its transposed ones
// sorted in registers, folded and compacted to the start of the row's combined
// segment in comb; olen = kept entries. Rows of more than 256 combined entries
// are written unsorted to their comb segment and listed in big[] for
// k_fuzzy_merge_mid.
__global__ __launch_bounds__(FZ_MB) void k_fz_bucket_merge(
    const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
    const float* __restrict__ w, uint32_t n, uint32_t shift, uint32_t nb,
    const uint32_t* __restrict__ bstart, const uint64_t* __restrict__ pay,
    const uint32_t* __restrict__ pj, uint64_t* __restrict__ srt_g, uint32_t* __restrict__ cnt_g,
    uint32_t* __restrict__ t_ptr, uint64_t* __restrict__ comb, uint32_t* __restrict__ olen,
    uint32_t* __restrict__ big, uint32_t* __restrict__ nbig) {
  extern __shared__ uint64_t lds[];
  uint64_t* s_srt = lds;
  uint64_t* s_wave = lds + FZ_BCAP;
  uint32_t* s_off = reinterpret_cast<uint32_t*>(s_wave + (FZ_MB / WAVE) * 4 * WAVE);
  uint32_t* s_cur = s_off + FZ_BROWS_LDS + 1;
  __shared__ uint32_t s_part[FZ_MB / WAVE];
  const uint32_t tid = threadIdx.x, l = tid & (WAVE - 1), wv = tid / WAVE;
  const uint32_t b = blockIdx.x;
  const uint32_t rows_b = 1u << shift, row0 = b << shift;
  const uint32_t rows = min(rows_b, n - row0);
  const uint32_t p0 = bstart[b], p1 = bstart[b + 1], c = p1 - p0;
  uint64_t* srt = c <= (uint32_t)FZ_BCAP ? s_srt : srt_g + p0;
  uint32_t* off = rows_b <= (uint32_t)FZ_BROWS_LDS ? s_off : cnt_g + (size_t)b * 2 * (rows_b + 1);
  uint32_t* cur = rows_b <= (uint32_t)FZ_BROWS_LDS ? s_cur : off + rows_b + 1;
  for (uint32_t r = tid; r < rows; r += FZ_MB) off[r] = 0;
  __syncthreads();
  for (uint32_t t = tid; t < c; t += FZ_MB) atomicAdd(&off[pj[p0 + t]], 1u);
  __syncthreads();
  fz_block_scan(off, rows, s_part);
  for (uint32_t r = tid; r < rows; r += FZ_MB) {
    cur[r] = off[r];
    t_ptr[row0 + r] = p0 + off[r];
  }
  if (b == nb - 1 && tid == 0) t_ptr[n] = p1;
  __syncthreads();
  for (uint32_t t = tid; t < c; t += FZ_MB) srt[atomicAdd(&cur[pj[p0 + t]], 1u)] = pay[p0 + t];
  __syncthreads();
  uint64_t* ws = s_wave + wv * 4 * WAVE;
  for (uint32_t r = wv; r < rows; r += FZ_MB / WAVE) {
    const uint32_t row = row0 + r;
    const uint32_t s = indptr[row], own = indptr[row + 1] - s;
    const uint32_t ts = off[r], len = own + (off[r + 1] - ts);
    const uint64_t coff = (uint64_t)s + p0 + ts;
    auto entry = [&](uint32_t q) -> uint64_t {
      return q < own ? fz_key(indices[s + q], 0, w[s + q]) : srt[ts + (q - own)];
    };
    if (len > 4 * WAVE) {
      for (uint32_t q = l; q < len; q += WAVE) comb[coff + q] = entry(q);
      if (l == 0) big[atomicAdd(nbig, 1u)] = row;
      continue;
    }
    if (len <= WAVE) {
      ws[l] = wave_sort64((uint32_t)l < len ? entry(l) : ~0ull);
    } else {
      uint64_t v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t e = q * WAVE + l;
        v[q] = e < len ? entry(e) : ~0ull;
      }
      wave_sort_n<4>(v);
#pragma unroll
      for (int q = 0; q < 4; ++q) ws[q * WAVE + l] = v[q];
    }
    fz_wave_sync();
    uint32_t kept = 0;
    for (uint32_t base = 0; base < len; base += WAVE) {
      const uint32_t e = base + l;
      bool keep = false;
      uint64_t outk = 0;
      if (e < len) {
        const uint32_t col = fz_col(ws[e]);
        if (e == 0 || fz_col(ws[e - 1]) != col) {
          float acc = 0.0f;
          for (uint32_t t = e; t < len && fz_col(ws[t]) == col; ++t) {
            const float x = fz_w(ws[t]);
            acc = acc + x - acc * x;
          }
          keep = acc > 1e-5f;
          outk = ((uint64_t)col << 32) | __float_as_uint(acc);
        }
      }
      const uint64_t mask = __ballot(keep);
      if (keep) comb[coff + kept + (uint32_t)__popcll(mask & ((1ull << l) - 1))] = outk;
      kept += (uint32_t)__popcll(mask);
    }
    if (l == 0) olen[row] = kept;
    fz_wave_sync();
  }
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
