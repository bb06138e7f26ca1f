// Learned seeding: a 50 -> 128 -> 32 MLP encoder (GELU), forward and backward
// written by hand for gfx950. Weights are kept in FP32 for the optimiser and
// rounded to BF16 for every forward and backward product; all sums are FP32.
//
//   h_pre = W1~ s(x) + b1,   h = gelu(h_pre),   z = W2~ h + b2
//   s(x)  = (x - mu) * inv_scale          (fixed input standardisation)
//   gelu(u) = u * Phi(u) = 0.5 u (1 + erf(u / sqrt 2))
//
// Triplet loss per (anchor, positive, negative):
//   L = max(0, |z_a - z_p|^2 - |z_a - z_n|^2 + margin),  averaged over the batch.
// Weight gradients are reduced in a fixed order (per-block partial sums, then
// one ordered sum), so training is deterministic for a seed.
//
// There is no reference counterpart (the reference has no learned component);
// the seeding it feeds replaces nndescent.rs:767-817's forest start.
#include "psq_kernels.hpp"

namespace psq {

constexpr int MLP_IN = 50, MLP_H = 128, MLP_OUT = 32;

__device__ inline float bf16_round(float v) { return (float)(__bf16)v; }

// erf is the device library's; Phi'(u) = exp(-u^2/2) / sqrt(2 pi).
__device__ inline float gelu(float u) { return 0.5f * u * (1.0f + erff(u * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float u) {
  const float cdf = 0.5f * (1.0f + erff(u * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * u * u);
  return cdf + u * pdf;
}

// Round the FP32 master weights to BF16 values (kept as float for the VALU).
__global__ void k_mlp_round(const float* __restrict__ w, float* __restrict__ wr, uint32_t count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) wr[i] = bf16_round(w[i]);
}

// Parameter block layout (floats): W1[128][50], b1[128], W2[32][128], b2[32].
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_OUT * MLP_H;
constexpr int N_PARAMS = OFF_B2 + MLP_OUT;

// Forward for `rows` points given by index (or 0..rows-1 when idx is NULL) of
// X (row stride ldx). One block of 128 threads per 16 rows: thread j computes
// hidden unit j for the 16 rows, then 32 threads compute the outputs.
// Writes h_pre (optional, for backward) and z.
__global__ __launch_bounds__(128) void k_mlp_forward(
    const float* __restrict__ X, uint32_t ldx, const uint32_t* __restrict__ idx, uint32_t rows,
    const float* __restrict__ pr, float mu, float inv_scale, float* __restrict__ hpre_out,
    float* __restrict__ z_out) {
  __shared__ float s_x[16][MLP_IN];
  __shared__ float s_h[16][MLP_H];
  const uint32_t r0 = blockIdx.x * 16;
  for (uint32_t e = threadIdx.x; e < 16 * MLP_IN; e += blockDim.x) {
    const uint32_t rr = e / MLP_IN, t = e % MLP_IN;
    const uint32_t r = r0 + rr;
    float v = 0.0f;
    if (r < rows) {
      const uint32_t src = idx ? idx[r] : r;
      v = (X[(size_t)src * ldx + t] - mu) * inv_scale;
    }
    s_x[rr][t] = v;
  }
  __syncthreads();
  const uint32_t j = threadIdx.x;  // hidden unit
  const float* w1 = pr + OFF_W1 + j * MLP_IN;
  float wreg[MLP_IN];
#pragma unroll
  for (int t = 0; t < MLP_IN; ++t) wreg[t] = w1[t];
  const float b1 = pr[OFF_B1 + j];
  for (int rr = 0; rr < 16; ++rr) {
    float acc = b1;
#pragma unroll
    for (int t = 0; t < MLP_IN; ++t) acc = fmaf(wreg[t], s_x[rr][t], acc);
    const uint32_t r = r0 + rr;
    if (hpre_out && r < rows) hpre_out[(size_t)r * MLP_H + j] = acc;
    s_h[rr][j] = gelu(acc);
  }
  __syncthreads();
  // outputs: 128 threads = 32 outputs x 4 row groups of 4 rows
  const uint32_t o = threadIdx.x % MLP_OUT, rg = threadIdx.x / MLP_OUT;
  const float* w2 = pr + OFF_W2 + o * MLP_H;
  const float b2 = pr[OFF_B2 + o];
  for (int q = 0; q < 4; ++q) {
    const uint32_t rr = rg * 4 + q;
    float acc = b2;
    for (int t = 0; t < MLP_H; ++t) acc = fmaf(w2[t], s_h[rr][t], acc);
    const uint32_t r = r0 + rr;
    if (r < rows) z_out[(size_t)r * MLP_OUT + o] = acc;
  }
}

// ── encode on the FP32 matrix cores ──
// z for every row of X (stride 50) with v_mfma_f32_16x16x4f32, 16 rows per
// wave-tile. MFMA 16x16x4 operands: lane l (c = l % 16, g = l / 16) holds
// A[c][g] and B[g][c]; D[4g + i][c] is its i-th output.
//   layer 1, transposed: H^T (hidden x row) = W1 x X^T, tile t of 16 hidden
//     units, 13 k-steps over the 50 inputs (52 with zeros): lane l holds
//     H[row c][hidden 16t + 4g + i] in D.
//   layer 2: Z = gelu(H) x W2^T with the hidden index permuted so that each
//     lane's A operand is one of its own layer-1 outputs: k-step s = 4t + i
//     contracts hidden 16t + 4g + i in lane group g (every hidden unit once).
// The per-lane weight operands are packed once (k_encode_pack) into
//   A1[t][s][l] = W1[16t + c][4s + g]   (t < 8, s < 13; 0 past input 50)
//   B2[u][s][l] = W2[16u + c][16(s / 4) + 4g + s % 4]   (u < 2, s < 32)
// and staged in LDS per block. Products are FP32; sums are ordered by the
// matrix core, so z agrees with the VALU forward to FP32 rounding.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int ENC_KS1 = 13, ENC_KS2 = 32;
constexpr int ENC_A1 = 8 * ENC_KS1 * 64, ENC_B2 = 2 * ENC_KS2 * 64;
static_assert(ENC_A1 + ENC_B2 == ENC_PACKED, "packed encoder operands");

__global__ void k_encode_pack(const float* __restrict__ pr, float* __restrict__ packed) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (uint32_t)(ENC_A1 + ENC_B2)) return;
  const uint32_t l = e % 64, c = l % 16, g = l / 16;
  if (e < (uint32_t)ENC_A1) {
    const uint32_t s = (e / 64) % ENC_KS1, t = e / (64 * ENC_KS1);
    const uint32_t in = 4 * s + g;
    packed[e] = in < (uint32_t)MLP_IN ? pr[OFF_W1 + (16 * t + c) * MLP_IN + in] : 0.0f;
  } else {
    const uint32_t f = e - ENC_A1;
    const uint32_t s = (f / 64) % ENC_KS2, u = f / (64 * ENC_KS2);
    const uint32_t h = 16 * (s / 4) + 4 * g + s % 4;
    packed[e] = pr[OFF_W2 + (16 * u + c) * MLP_H + h];
  }
}

__global__ __launch_bounds__(ENC_WAVES * 64) void k_encode_mfma(
    const float* __restrict__ X, uint32_t rows, const float* __restrict__ packed,
    const float* __restrict__ pr, float mu, float inv_scale, float* __restrict__ z_out) {
  __shared__ float s_w[ENC_A1 + ENC_B2];
  __shared__ float s_b[MLP_H + MLP_OUT];
  for (uint32_t e = threadIdx.x; e < (uint32_t)(ENC_A1 + ENC_B2); e += blockDim.x) s_w[e] = packed[e];
  for (uint32_t e = threadIdx.x; e < (uint32_t)(MLP_H + MLP_OUT); e += blockDim.x)
    s_b[e] = pr[e < (uint32_t)MLP_H ? OFF_B1 + e : OFF_B2 + (e - MLP_H)];
  __syncthreads();
  const uint32_t l = threadIdx.x % 64, c = l % 16, g = l / 16;
  const uint32_t ntiles = (rows + 15) / 16;
  const uint32_t stride = gridDim.x * ENC_WAVES;
  for (uint32_t tile = blockIdx.x * ENC_WAVES + threadIdx.x / 64; tile < ntiles; tile += stride) {
    // keeps the weight reads from LDS inside the loop: hoisted, they would take
    // 168 registers per lane and leave one wave per SIMD
    asm volatile("" ::: "memory");
    const uint32_t r0 = tile * 16;
    float xb[ENC_KS1];
    const bool rv = r0 + c < rows;
#pragma unroll
    for (int s = 0; s < ENC_KS1; ++s) {
      const uint32_t in = 4 * s + g;
      xb[s] = (rv && in < (uint32_t)MLP_IN) ? (X[(size_t)(r0 + c) * MLP_IN + in] - mu) * inv_scale : 0.0f;
    }
    float h[8][4];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      f32x4 acc;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = s_b[16 * t + 4 * g + i];
#pragma unroll
      for (int s = 0; s < ENC_KS1; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(s_w[(t * ENC_KS1 + s) * 64 + l], xb[s], acc, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) h[t][i] = gelu(acc[i]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      f32x4 acc;
      const float b2 = s_b[MLP_H + 16 * u + c];
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = b2;
#pragma unroll
      for (int s = 0; s < ENC_KS2; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(h[s / 4][s % 4],
                                                   s_w[ENC_A1 + (u * ENC_KS2 + s) * 64 + l], acc, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t r = r0 + 4 * g + i;
        if (r < rows) z_out[(size_t)r * MLP_OUT + 16 * u + c] = acc[i];
      }
    }
  }
}

// Triplet loss and dL/dz for a batch of B triplets laid out as rows
// [a_0..a_{B-1}, p_0..p_{B-1}, n_0..n_{B-1}] of z. One thread per triplet.
__global__ void k_triplet_loss(const float* __restrict__ z, uint32_t B, float margin,
                               float* __restrict__ dz, float* __restrict__ loss_terms) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float* za = z + (size_t)b * MLP_OUT;
  const float* zp = z + (size_t)(B + b) * MLP_OUT;
  const float* zn = z + (size_t)(2 * B + b) * MLP_OUT;
  float dap = 0.0f, dan = 0.0f;
  for (int t = 0; t < MLP_OUT; ++t) {
    const float u = za[t] - zp[t], v = za[t] - zn[t];
    dap = fmaf(u, u, dap);
    dan = fmaf(v, v, dan);
  }
  const float l = dap - dan + margin;
  const bool active = l > 0.0f;
  loss_terms[b] = active ? l : 0.0f;
  const float s = active ? 2.0f / (float)B : 0.0f;
  float* ga = dz + (size_t)b * MLP_OUT;
  float* gp = dz + (size_t)(B + b) * MLP_OUT;
  float* gn = dz + (size_t)(2 * B + b) * MLP_OUT;
  for (int t = 0; t < MLP_OUT; ++t) {
    ga[t] = s * (zn[t] - zp[t]);
    gp[t] = -s * (za[t] - zp[t]);
    gn[t] = s * (za[t] - zn[t]);
  }
}

// Backward, per block of 16 rows: partial weight gradients for the block.
// partial layout per block: N_PARAMS floats. Thread j handles hidden unit j.
__global__ __launch_bounds__(128) void k_mlp_backward(
    const float* __restrict__ X, uint32_t ldx, const uint32_t* __restrict__ idx, uint32_t rows,
    const float* __restrict__ pr, float mu, float inv_scale, const float* __restrict__ hpre,
    const float* __restrict__ dz, float* __restrict__ partial) {
  __shared__ float s_x[16][MLP_IN];
  __shared__ float s_dz[16][MLP_OUT];
  __shared__ float s_dpre[16][MLP_H];
  const uint32_t r0 = blockIdx.x * 16;
  for (uint32_t e = threadIdx.x; e < 16 * MLP_IN; e += blockDim.x) {
    const uint32_t rr = e / MLP_IN, t = e % MLP_IN;
    const uint32_t r = r0 + rr;
    float v = 0.0f;
    if (r < rows) {
      const uint32_t src = idx ? idx[r] : r;
      v = (X[(size_t)src * ldx + t] - mu) * inv_scale;
    }
    s_x[rr][t] = v;
  }
  for (uint32_t e = threadIdx.x; e < 16 * MLP_OUT; e += blockDim.x) {
    const uint32_t rr = e / MLP_OUT, t = e % MLP_OUT;
    const uint32_t r = r0 + rr;
    s_dz[rr][t] = r < rows ? dz[(size_t)r * MLP_OUT + t] : 0.0f;
  }
  __syncthreads();
  float* out = partial + (size_t)blockIdx.x * N_PARAMS;
  const uint32_t j = threadIdx.x;
  // dh_j = sum_o W2[o][j] dz_o ; dpre_j = dh_j * gelu'(hpre_j)
  // dW2[o][j] = sum_r dz[r][o] * gelu(hpre[r][j])
  float dw2[MLP_OUT];
#pragma unroll
  for (int o = 0; o < MLP_OUT; ++o) dw2[o] = 0.0f;
  float db1 = 0.0f;
  for (int rr = 0; rr < 16; ++rr) {
    const uint32_t r = r0 + rr;
    const float hp = r < rows ? hpre[(size_t)r * MLP_H + j] : 0.0f;
    const float h = r < rows ? gelu(hp) : 0.0f;
    float dh = 0.0f;
#pragma unroll
    for (int o = 0; o < MLP_OUT; ++o) {
      dh = fmaf(pr[OFF_W2 + o * MLP_H + j], s_dz[rr][o], dh);
      dw2[o] = fmaf(s_dz[rr][o], h, dw2[o]);
    }
    const float dp = r < rows ? dh * gelu_grad(hp) : 0.0f;
    s_dpre[rr][j] = dp;
    db1 += dp;
  }
#pragma unroll
  for (int o = 0; o < MLP_OUT; ++o) out[OFF_W2 + o * MLP_H + j] = dw2[o];
  out[OFF_B1 + j] = db1;
  __syncthreads();
  // dW1[j][t] = sum_r dpre[r][j] * x[r][t]
  for (int t = 0; t < MLP_IN; ++t) {
    float acc = 0.0f;
    for (int rr = 0; rr < 16; ++rr) acc = fmaf(s_dpre[rr][j], s_x[rr][t], acc);
    out[OFF_W1 + j * MLP_IN + t] = acc;
  }
  if (j < MLP_OUT) {
    float acc = 0.0f;
    for (int rr = 0; rr < 16; ++rr) acc += s_dz[rr][j];
    out[OFF_B2 + j] = acc;
  }
}

// Ordered sum of the per-block partials: grad[p] = sum_b partial[b][p].
__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t nblocks,
                             float* __restrict__ grad) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)N_PARAMS) return;
  float s = 0.0f;
  for (uint32_t b = 0; b < nblocks; ++b) s += partial[(size_t)b * N_PARAMS + p];
  grad[p] = s;
}

// Adam step on the FP32 master weights, then BF16 rounding for the next pass.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ wr, const float* __restrict__ g,
                       float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)N_PARAMS) return;
  const float gg = g[p];
  const float a = b1 * m1[p] + (1.0f - b1) * gg;
  const float v = b2 * m2[p] + (1.0f - b2) * gg * gg;
  m1[p] = a;
  m2[p] = v;
  const float mh = a / c1, vh = v / c2;
  const float nw = w[p] - lr * mh / (sqrtf(vh) + eps);
  w[p] = nw;
  wr[p] = bf16_round(nw);
}

// Sum and sum of squares of all entries, per block of `chunk` rows, summed in
// a fixed order (thread-strided then a fixed tree), for the standardisation.
__global__ __launch_bounds__(256) void k_moments(const float* __restrict__ X, uint32_t n, uint32_t d,
                                                 uint32_t chunk, double* __restrict__ partial) {
  __shared__ double s1[256], s2[256];
  const uint32_t r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
  double a = 0.0, b = 0.0;
  for (size_t e = (size_t)r0 * d + threadIdx.x; e < (size_t)r1 * d; e += blockDim.x) {
    const double v = X[e];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s1[0];
    partial[2 * blockIdx.x + 1] = s2[0];
  }
}

// dst[r][t] = src[idx[r]][t] for r < rows, t < d.
__global__ void k_gather_rows(const float* __restrict__ src, uint32_t d, const uint32_t* __restrict__ idx,
                              uint32_t rows, float* __restrict__ dst) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)rows * d) return;
  const uint32_t r = (uint32_t)(e / d), t = (uint32_t)(e - (size_t)r * d);
  dst[e] = src[(size_t)idx[r] * d + t];
}

// Triplet sampling for one step: anchors uniform over the subset, positive
// uniform over the anchor's exact top-kk, negative uniform over the subset.
// Rows written as [anchors | positives | negatives] (indices into the subset).
__global__ void k_triplets(const uint32_t* __restrict__ knn, uint32_t kk, uint32_t ns, uint32_t B,
                           uint64_t seed, uint64_t step, uint32_t* __restrict__ rows) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const uint64_t key = step * 0x100000000ull + b;
  const uint32_t a = (uint32_t)hash_index(seed, key, 1, ns);
  const uint32_t pi = (uint32_t)hash_index(seed, key, 2, kk);
  const uint32_t ng = (uint32_t)hash_index(seed, key, 3, ns);
  rows[b] = a;
  rows[B + b] = knn[(size_t)a * kk + pi];
  rows[2 * B + b] = ng;
}

}  // namespace psq
