// Row scaling and distance conversion of pp.neighbors' cosine metric on the
// approximate path (python/polariseq/__init__.py:2540-2552):
//   search = emb / max(np.linalg.norm(emb, axis=1), 1e-12)        (float32)
//   dist   = float32(float64(dist) ** 2 / 2)   (|a - b|^2 = 2 - 2 cos on unit rows)
// Both bit for bit as numpy computes them: the squared norm is numpy's
// pairwise float32 sum of x * x over the row (8 running sums seeded with the
// first 8 elements, combined ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)),
// the rest added in order; rows of more than 128 split at n/2 rounded down to
// a multiple of 8, recursively), the square root and the division correctly
// rounded (the f32 square root through f64, which is exact after rounding
// since 53 >= 2 * 24 + 2; f32 division is IEEE in HIP's default mode).
#include "psq_kernels.hpp"

namespace psq {

// numpy's pairwise sum of x[i]^2 for n <= 128 (float32 throughout)
__device__ inline float pw_sumsq_block(const float* __restrict__ x, uint32_t n) {
  if (n < 8) {
    float res = 0.0f;
    for (uint32_t i = 0; i < n; ++i) res = res + x[i] * x[i];
    return res;
  }
  float r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = x[j] * x[j];
  uint32_t i = 8;
  for (; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = r[j] + x[i + j] * x[i + j];
  }
  float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res = res + x[i] * x[i];
  return res;
}

// numpy's split of a block longer than 128: n / 2 rounded down to a multiple of 8.
__device__ inline uint32_t pw_split(uint32_t n) { return n / 2 - (n / 2) % 8; }
// The same for n <= ROWS_D_MAX = 256. The left part of a split is <= 128; the
// right part is at most n / 2 + 7 <= 135 and is split once more when > 128.
__device__ inline float pw_sumsq(const float* __restrict__ x, uint32_t n) {
  if (n <= 128) return pw_sumsq_block(x, n);
  const uint32_t n2 = pw_split(n), m = n - n2;
  const float left = pw_sumsq_block(x, n2);
  if (m <= 128) return left + pw_sumsq_block(x + n2, m);
  const uint32_t m2 = pw_split(m);
  return left + (pw_sumsq_block(x + n2, m2) + pw_sumsq_block(x + n2 + m2, m - m2));
}

// One thread per row: out = x / max(|x|, 1e-12f).
__global__ void k_unit_rows(const float* __restrict__ x, uint64_t n, uint32_t d,
                            float* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* xi = x + i * d;
  const float s = pw_sumsq(xi, d);
  float nrm = (float)__dsqrt_rn((double)s);
  nrm = nrm > 1e-12f ? nrm : 1e-12f;  // np.maximum(norms, 1e-12) in float32
  float* oi = out + i * d;
  for (uint32_t t = 0; t < d; ++t) oi[t] = xi[t] / nrm;
}

// dist = float32(float64(dist)^2 / 2): the square of a float32 is exact in
// float64 and the halving is exact, so one rounding, as numpy's.
__global__ void k_half_square(float* __restrict__ dist, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double v = (double)dist[i];
  dist[i] = (float)(v * v / 2.0);
}

}  // namespace psq
