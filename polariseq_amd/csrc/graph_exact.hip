// The exact neighbour graph of small embeddings: neighbors() / knn_for_one of
// crates/polariseq-core/src/neighbors.rs:101-241, the path pp.neighbors takes
// up to 10,000 cells.
//
// The reference keeps, per query, a bounded max-heap of k = min(n_neighbors,
// n) entries while it scans every point in index order (strict `<` against the
// root once the heap is full), then sorts the heap array stably by distance
// and drops entry 0 as self. Which of several equal distances survive at the
// boundary, the order of equal distances in the row, and whether entry 0 is
// the query itself or a duplicate of it all depend on the heap's array layout,
// so the scan below runs that heap, step for step, one lane per query:
//
// k_nbr_heap    one lane per query (64-lane blocks), every point in index
//               order. The distance is the reference's f32 value: the 8-lane
//               sq_dist (simd.rs:45-62), or cosine 1 - dot/(sqrt(na) sqrt(nb))
//               with 1.0 when the denominator is 0 (neighbors.rs:232-241).
//               Point rows are read at uniform addresses (all lanes of a wave
//               read the same row). The heap lives in LDS (k <= NBR_LDS_K) or
//               in the global scratch rows; it ends in the global rows,
//               slot-major ([slot][n]).
// k_nbr_sort*   sorts each heap by (distance, array position), the order of a
//               stable sort on distance, and writes entries 1 .. k-1 (square
//               roots for Euclidean).
#include "psq_kernels.hpp"

namespace psq {

// simd.rs:22-37 with the dimension fixed at compile time (operand in registers).
template <int D, typename PA, typename PB>
__device__ inline float dot_fixed(PA p, PB q) {
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ch = 0; ch < D / 8; ++ch) {
#pragma unroll
    for (int l = 0; l < 8; ++l) acc[l] = acc[l] + p[ch * 8 + l] * q[ch * 8 + l];
  }
  float tail = 0.f;
#pragma unroll
  for (int i = (D / 8) * 8; i < D; ++i) tail = tail + p[i] * q[i];
  return (((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]))) + tail;
}

// Correctly rounded f32 square root and quotient (Rust's f32 sqrt and `/`):
// computed in FP64 and rounded once to FP32, exact after rounding because
// 53 >= 2 * 24 + 2.
__device__ inline float sqrt_rn(float v) { return (float)__dsqrt_rn((double)v); }
__device__ inline float div_rn(float a, float b) { return (float)((double)a / (double)b); }

// rs[i] = sqrt(dot(x_i, x_i)) in f32, the factors of cosine_dist's denominator.
__global__ void k_nbr_row_norms(const float* __restrict__ X, uint32_t n, uint32_t d,
                                float* __restrict__ rs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* x = X + (size_t)i * d;
  rs[i] = sqrt_rn(dot_ref(x, x, (int)d));
}

// Distance of the reference's knn_for_one, neighbors.rs:189-192, 228-241.
template <int D, int M>
__device__ inline float nbr_dist(const float* qv, const float* __restrict__ o, uint32_t d, float rq,
                                 float rj) {
  if (M == 0) return D > 0 ? sq_dist_fixed<D>(qv, o) : sq_dist_ref(qv, o, (int)d);
  const float dt = D > 0 ? dot_fixed<D>(qv, o) : dot_ref(qv, o, (int)d);
  const float den = rq * rj;
  return den > 0.f ? 1.0f - div_rn(dt, den) : 1.0f;
}

// Slot-major heap view of one lane: slot s at h[s * stride].
struct LaneHeap {
  float* d;
  uint32_t* i;
  uint32_t stride;
  __device__ float dist(uint32_t s) const { return d[(size_t)s * stride]; }
  __device__ void set(uint32_t s, float v, uint32_t j) {
    d[(size_t)s * stride] = v;
    i[(size_t)s * stride] = j;
  }
  __device__ void swap(uint32_t a, uint32_t b) {
    const float da = dist(a), db = dist(b);
    const uint32_t ia = i[(size_t)a * stride], ib = i[(size_t)b * stride];
    set(a, db, ib);
    set(b, da, ia);
  }
};

template <int D, int M>
__global__ __launch_bounds__(64) void k_nbr_heap(const float* __restrict__ X, uint32_t n, uint32_t d,
                                                 uint32_t k, const float* __restrict__ rs,
                                                 float* __restrict__ hd, uint32_t* __restrict__ hi,
                                                 uint32_t lds_heap) {
  extern __shared__ uint8_t nbr_smem[];
  const uint32_t lane = threadIdx.x;
  const uint32_t q = blockIdx.x * 64 + lane;
  if (q >= n) return;
  const float* xq = X + (size_t)q * d;
  float qbuf[D > 0 ? D : 1];
  const float* qv = xq;
  if (D > 0) {
#pragma unroll
    for (int t = 0; t < (D > 0 ? D : 1); ++t) qbuf[t] = xq[t];
    qv = qbuf;
  }
  const float rq = M == 1 ? rs[q] : 0.f;
  LaneHeap h;
  if (lds_heap) {
    h.d = reinterpret_cast<float*>(nbr_smem) + lane;
    h.i = reinterpret_cast<uint32_t*>(nbr_smem + (size_t)k * 64 * sizeof(float)) + lane;
    h.stride = 64;
  } else {
    h.d = hd + q;
    h.i = hi + q;
    h.stride = n;
  }
  // The first k points are pushed (the heap is not full), each sifted up.
  for (uint32_t j = 0; j < k; ++j) {
    const float v = nbr_dist<D, M>(qv, X + (size_t)j * d, d, rq, M == 1 ? rs[j] : 0.f);
    h.set(j, v, j);
    uint32_t c = j;
    while (c > 0) {
      const uint32_t p = (c - 1) / 2;
      if (h.dist(c) > h.dist(p)) {
        h.swap(c, p);
        c = p;
      } else {
        break;
      }
    }
  }
  // Then a point replaces the root when strictly closer, and is sifted down.
  float root = h.dist(0);
  for (uint32_t j = k; j < n; ++j) {
    const float v = nbr_dist<D, M>(qv, X + (size_t)j * d, d, rq, M == 1 ? rs[j] : 0.f);
    if (v < root) {
      h.set(0, v, j);
      uint32_t p = 0;
      for (;;) {
        const uint32_t l = 2 * p + 1, r = 2 * p + 2;
        uint32_t m = p;
        if (l < k && h.dist(l) > h.dist(m)) m = l;
        if (r < k && h.dist(r) > h.dist(m)) m = r;
        if (m == p) break;
        h.swap(p, m);
        p = m;
      }
      root = h.dist(0);
    }
  }
  if (lds_heap)
    for (uint32_t s = 0; s < k; ++s) {
      hd[(size_t)s * n + q] = h.dist(s);
      hi[(size_t)s * n + q] = h.i[(size_t)s * 64];
    }
}

// Sort key of heap slot s: an order-preserving image of the f32 distance
// (negative values can occur for cosine; -0 counts as +0, as partial_cmp
// says) above the slot, so equal distances keep their array order, the
// order of Rust's stable sort_by. A NaN (non-finite input) sorts last.
__device__ inline uint64_t nbr_key(float v, uint32_t s) {
  uint32_t b = __float_as_uint(v);
  if (b == 0x80000000u) b = 0;
  b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return ((uint64_t)b << 32) | s;
}

__device__ inline void nbr_emit(const float* __restrict__ hd, const uint32_t* __restrict__ hi,
                                uint32_t n, uint32_t q, uint32_t k, int metric, uint32_t rank,
                                uint64_t key, uint32_t* __restrict__ out_idx,
                                float* __restrict__ out_d) {
  const uint32_t s = (uint32_t)key;
  const float v = hd[(size_t)s * n + q];
  const size_t o = (size_t)q * (k - 1) + (rank - 1);
  out_idx[o] = hi[(size_t)s * n + q];
  out_d[o] = metric == 0 ? sqrt_rn(v) : v;
}

// k <= 64: one wave per query.
__global__ __launch_bounds__(64) void k_nbr_sort64(const float* __restrict__ hd,
                                                   const uint32_t* __restrict__ hi, uint32_t n,
                                                   uint32_t k, int metric,
                                                   uint32_t* __restrict__ out_idx,
                                                   float* __restrict__ out_d) {
  const uint32_t q = block_id(), l = threadIdx.x;
  if (q >= n) return;  // block-uniform
  uint64_t key = l < k ? nbr_key(hd[(size_t)l * n + q], l) : EMPTY_KEY;
  key = wave_sort64(key);
  if (l >= 1 && l < k) nbr_emit(hd, hi, n, q, k, metric, l, key, out_idx, out_d);
}

// k > 64: one block per query, a bitonic sort of P (power of two >= k) keys in LDS.
__global__ __launch_bounds__(256) void k_nbr_sort_lds(const float* __restrict__ hd,
                                                      const uint32_t* __restrict__ hi, uint32_t n,
                                                      uint32_t k, uint32_t P, int metric,
                                                      uint32_t* __restrict__ out_idx,
                                                      float* __restrict__ out_d) {
  extern __shared__ uint64_t nbr_keys[];
  const uint32_t q = block_id();
  if (q >= n) return;  // block-uniform
  for (uint32_t s = threadIdx.x; s < P; s += blockDim.x)
    nbr_keys[s] = s < k ? nbr_key(hd[(size_t)s * n + q], s) : EMPTY_KEY;
  __syncthreads();
  for (uint32_t size = 2; size <= P; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = threadIdx.x; t < P / 2; t += blockDim.x) {
        const uint32_t a = 2 * t - (t & (stride - 1)), b = a + stride;
        const bool up = (a & size) == 0;
        const uint64_t x = nbr_keys[a], y = nbr_keys[b];
        if ((x > y) == up) {
          nbr_keys[a] = y;
          nbr_keys[b] = x;
        }
      }
      __syncthreads();
    }
  for (uint32_t r = 1 + threadIdx.x; r < k; r += blockDim.x)
    nbr_emit(hd, hi, n, q, k, metric, r, nbr_keys[r], out_idx, out_d);
}

// Row pointers of a graph with `per` entries in every row.
__global__ void k_nbr_indptr(uint32_t* __restrict__ indptr, uint32_t n, uint32_t per) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) indptr[i] = i * per;
}

template __global__ void k_nbr_heap<50, 0>(const float*, uint32_t, uint32_t, uint32_t, const float*,
                                           float*, uint32_t*, uint32_t);
template __global__ void k_nbr_heap<50, 1>(const float*, uint32_t, uint32_t, uint32_t, const float*,
                                           float*, uint32_t*, uint32_t);
template __global__ void k_nbr_heap<30, 0>(const float*, uint32_t, uint32_t, uint32_t, const float*,
                                           float*, uint32_t*, uint32_t);
template __global__ void k_nbr_heap<30, 1>(const float*, uint32_t, uint32_t, uint32_t, const float*,
                                           float*, uint32_t*, uint32_t);
template __global__ void k_nbr_heap<0, 0>(const float*, uint32_t, uint32_t, uint32_t, const float*,
                                          float*, uint32_t*, uint32_t);
template __global__ void k_nbr_heap<0, 1>(const float*, uint32_t, uint32_t, uint32_t, const float*,
                                          float*, uint32_t*, uint32_t);

}  // namespace psq
