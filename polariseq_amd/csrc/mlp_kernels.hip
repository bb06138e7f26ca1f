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

// ── one training step in two launches ──
// k_mlp_step: a block of MLP_STEP_T triplets (3 T rows: RT = 3 T / 16 16-row
// tiles), forward, triplet loss, backward, its partial gradient (N_PARAMS
// floats per block). Triplets come from `rows` ([anchors | positives |
// negatives] of the whole batch) or, when rows is NULL, are sampled here:
// anchor = pool[hash % A] (cells whose exact top-kk is in knn), positive
// uniform over that top-kk, negative uniform over the subset.
// The five products run on the FP32 matrix cores (v_mfma_f32_16x16x4f32, the
// operand layout of k_encode_mfma: lane l, c = l % 16, g = l / 16, holds A[c][g]
// and B[g][c], and D[4g + i][c] is its i-th output), with FP32 activations and
// the BF16-valued weights:
//   H_pre = X W1^T + b1      (3T x 128, k = 52)      Z = gelu(H_pre) W2^T + b2 (3T x 32, k = 128)
//   dW2 = dZ^T gelu(H_pre)   (32 x 128, k = 3T)      dH = dZ W2              (3T x 128, k = 32)
//   dW1 = dH_pre^T X         (128 x 52, k = 3T),     dH_pre = dH * gelu'(H_pre)
// Activations live in LDS at row strides of 4 (x, h) and 52 (z) mod 64 words,
// so the 16 rows of a tile read by c fall on distinct banks: 144 KB for 96 rows,
// one block of 16 waves per CU. The block's STEP_WAVES waves take the tiles of
// each product in turn. The partial gradients are stored by tiles of 32 parameters, the blocks'
// slices of a tile side by side (pslot), so that the reduction streams them.
// k_mlp_reduce_adam: sums the blocks' partials in a fixed order and takes the
// Adam step (deterministic for a seed).
constexpr int STEP_R = 3 * MLP_STEP_T;  // rows per block
constexpr int RT = STEP_R / 16;         // 16-row tiles per block
static_assert(STEP_R % 16 == 0, "the block's rows are whole 16-row tiles");
constexpr int SX = 68, SH = 132, SZ = 52;  // LDS row strides (floats)
static_assert(N_PARAMS % 32 == 0, "the partials are stored by tiles of 32 parameters");

// Where block blk of nb puts its partial of parameter p.
__device__ inline size_t pslot(uint32_t p, uint32_t blk, uint32_t nb) {
  return ((size_t)(p / 32) * nb + blk) * 32 + p % 32;
}
constexpr int STEP_WAVES = MLP_STEP_WAVES;
constexpr size_t STEP_SMEM = (size_t)STEP_R * (SX + 2 * SH + SZ) * sizeof(float) + STEP_R * 4;

__device__ inline uint32_t step_row(const uint32_t* __restrict__ rows, uint32_t B, uint32_t b,
                                    uint32_t role, const uint32_t* __restrict__ pool, uint32_t A,
                                    const uint32_t* __restrict__ knn, uint32_t kk, uint32_t ns,
                                    uint64_t seed, uint64_t step) {
  if (rows) return rows[role * B + b];
  const uint64_t key = step * 0x100000000ull + b;
  const uint32_t a = (uint32_t)hash_index(seed, key, 1, A);
  if (role == 0) return pool[a];
  if (role == 1) return knn[(size_t)a * kk + (uint32_t)hash_index(seed, key, 2, kk)];
  return (uint32_t)hash_index(seed, key, 3, ns);
}

__global__ __launch_bounds__(64 * STEP_WAVES) void k_mlp_step(
    const float* __restrict__ X, const uint32_t* __restrict__ rows, uint32_t B,
    const uint32_t* __restrict__ pool, uint32_t A, const uint32_t* __restrict__ knn, uint32_t kk,
    uint32_t ns, uint64_t seed, uint64_t step, const float* __restrict__ pr, float mu,
    float inv_scale, float margin, float* __restrict__ partial, float* __restrict__ loss_terms) {
  extern __shared__ float step_lds[];
  float* s_x = step_lds;                  // [3T][SX]: x standardised, columns 50..63 zero
  float* s_hp = s_x + STEP_R * SX;        // [3T][SH]: h_pre, then dL/dh_pre
  float* s_h = s_hp + STEP_R * SH;        // [3T][SH]: gelu(h_pre)
  float* s_z = s_h + STEP_R * SH;         // [3T][SZ]: z, then dL/dz
  uint32_t* s_row = reinterpret_cast<uint32_t*>(s_z + STEP_R * SZ);
  const uint32_t tid = threadIdx.x, l = tid % 64, w = tid / 64, c = l % 16, g = l / 16;
  const uint32_t b0 = blockIdx.x * MLP_STEP_T;
  if (tid < (uint32_t)STEP_R) {
    const uint32_t role = tid / MLP_STEP_T, b = b0 + tid % MLP_STEP_T;
    s_row[tid] = b < B ? step_row(rows, B, b, role, pool, A, knn, kk, ns, seed, step) : 0xFFFFFFFFu;
  }
  __syncthreads();
  for (uint32_t e = tid; e < (uint32_t)(STEP_R * 64); e += blockDim.x) {
    const uint32_t rr = e / 64, t = e % 64;
    const uint32_t src = s_row[rr];
    s_x[rr * SX + t] =
        (t < (uint32_t)MLP_IN && src != 0xFFFFFFFFu) ? (X[(size_t)src * MLP_IN + t] - mu) * inv_scale : 0.0f;
  }
  __syncthreads();
  const float* W1 = pr + OFF_W1;
  const float* W2 = pr + OFF_W2;
  // layer 1: 3 x 8 tiles of (16 rows, 16 hidden units)
  for (uint32_t tile = w; tile < RT * 8; tile += STEP_WAVES) {
    const uint32_t rt = tile / 8, ht = tile % 8, hj = 16 * ht + c;
    f32x4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = pr[OFF_B1 + hj];
#pragma unroll
    for (int s = 0; s < ENC_KS1; ++s) {
      const uint32_t t = 4 * s + g;
      const float wv = t < (uint32_t)MLP_IN ? W1[hj * MLP_IN + t] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(s_x[(16 * rt + c) * SX + t], wv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t r = 16 * rt + 4 * g + i;
      s_hp[r * SH + hj] = acc[i];
      s_h[r * SH + hj] = gelu(acc[i]);
    }
  }
  __syncthreads();
  // layer 2: 3 x 2 tiles of (16 rows, 16 outputs)
  for (uint32_t tile = w; tile < RT * 2; tile += STEP_WAVES) {
    const uint32_t rt = tile / 2, ot = tile % 2, oc = 16 * ot + c;
    f32x4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = pr[OFF_B2 + oc];
#pragma unroll 8
    for (int s = 0; s < MLP_H / 4; ++s)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(s_h[(16 * rt + c) * SH + 4 * s + g],
                                                 W2[oc * MLP_H + 4 * s + g], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) s_z[(16 * rt + 4 * g + i) * SZ + oc] = acc[i];
  }
  __syncthreads();
  // loss and dL/dz: 4 lanes per triplet (8 outputs each), written over z
  {
    float dap = 0.0f, dan = 0.0f;
    const uint32_t tt = tid / 4, part = tid % 4;
    const bool mine = tid < 4u * MLP_STEP_T;
    const bool live = mine && b0 + tt < B;
    float za[8], zp[8], zn[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      za[q] = mine ? s_z[tt * SZ + part * 8 + q] : 0.0f;
      zp[q] = mine ? s_z[(MLP_STEP_T + tt) * SZ + part * 8 + q] : 0.0f;
      zn[q] = mine ? s_z[(2 * MLP_STEP_T + tt) * SZ + part * 8 + q] : 0.0f;
      const float u = za[q] - zp[q], v = za[q] - zn[q];
      dap = fmaf(u, u, dap);
      dan = fmaf(v, v, dan);
    }
    // the 4 lanes of a triplet are consecutive: combine in a fixed order
    dap += __shfl_xor(dap, 1);
    dan += __shfl_xor(dan, 1);
    dap += __shfl_xor(dap, 2);
    dan += __shfl_xor(dan, 2);
    __syncthreads();  // every z read before dz overwrites it
    if (mine) {
      const float lv = dap - dan + margin;
      const bool active = live && lv > 0.0f;
      if (live && part == 0) loss_terms[b0 + tt] = active ? lv : 0.0f;
      const float sc = active ? 2.0f / (float)B : 0.0f;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        s_z[tt * SZ + part * 8 + q] = sc * (zn[q] - zp[q]);
        s_z[(MLP_STEP_T + tt) * SZ + part * 8 + q] = -sc * (za[q] - zp[q]);
        s_z[(2 * MLP_STEP_T + tt) * SZ + part * 8 + q] = sc * (za[q] - zn[q]);
      }
    }
  }
  __syncthreads();
  auto put = [&](uint32_t p, float v) { partial[pslot(p, blockIdx.x, gridDim.x)] = v; };
  // dW2 = dZ^T gelu(H_pre): 2 x 8 tiles of (16 outputs, 16 hidden), k over the block's rows
  for (uint32_t tile = w; tile < 16; tile += STEP_WAVES) {
    const uint32_t ot = tile / 8, jt = tile % 8;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < STEP_R / 4; ++s) {
      const uint32_t r = 4 * s + g;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(s_z[r * SZ + 16 * ot + c], s_h[r * SH + 16 * jt + c], acc,
                                                 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) put(OFF_W2 + (16 * ot + 4 * g + i) * MLP_H + 16 * jt + c, acc[i]);
  }
  // dH = dZ W2, then dL/dh_pre = dH * gelu'(h_pre) in place of h_pre: 3 x 8 tiles, k over 32 outputs
  for (uint32_t tile = w; tile < RT * 8; tile += STEP_WAVES) {
    const uint32_t rt = tile / 8, jt = tile % 8, hj = 16 * jt + c;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < MLP_OUT / 4; ++s)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(s_z[(16 * rt + c) * SZ + 4 * s + g],
                                                 W2[(4 * s + g) * MLP_H + hj], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float* hp = s_hp + (16 * rt + 4 * g + i) * SH + hj;
      *hp = acc[i] * gelu_grad(*hp);
    }
  }
  __syncthreads();
  // biases: column sums over the block's rows, in row order
  if (tid < (uint32_t)MLP_H) {
    float acc = 0.0f;
    for (int r = 0; r < STEP_R; ++r) acc += s_hp[r * SH + tid];
    put(OFF_B1 + tid, acc);
  } else if (tid < (uint32_t)(MLP_H + MLP_OUT)) {
    const uint32_t o = tid - MLP_H;
    float acc = 0.0f;
    for (int r = 0; r < STEP_R; ++r) acc += s_z[r * SZ + o];
    put(OFF_B2 + o, acc);
  }
  // dW1 = dH_pre^T X: 8 x 4 tiles of (16 hidden, 16 inputs), k over the block's rows
  for (uint32_t tile = w; tile < 32; tile += STEP_WAVES) {
    const uint32_t jt = tile / 4, tt = tile % 4;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < STEP_R / 4; ++s) {
      const uint32_t r = 4 * s + g;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(s_hp[r * SH + 16 * jt + c], s_x[r * SX + 16 * tt + c], acc,
                                                 0, 0, 0);
    }
    const uint32_t t = 16 * tt + c;
    if (t < (uint32_t)MLP_IN) {
#pragma unroll
      for (int i = 0; i < 4; ++i) put(OFF_W1 + (16 * jt + 4 * g + i) * MLP_IN + t, acc[i]);
    }
  }
}

size_t mlp_step_smem() { return STEP_SMEM; }

// Block of 1024 threads = 32 parameters x 32 groups: group g sums the partials
// of blocks g, g + 32, ... into 4 accumulators by turn (block g + 32 i into
// accumulator i % 4, four loads in flight), added in a fixed tree, then the 32
// group sums are added in a fixed tree.
__global__ __launch_bounds__(1024) void k_mlp_reduce_adam(
    const float* __restrict__ partial, uint32_t nblocks, float* __restrict__ w,
    float* __restrict__ wr, float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
    float b2, float eps, float c1, float c2, float* __restrict__ grad_out) {
  __shared__ float s_g[32][33];
  const uint32_t pl = threadIdx.x % 32, grp = threadIdx.x / 32;
  const uint32_t p = blockIdx.x * 32 + pl;
  float q[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (p < (uint32_t)N_PARAMS) {
    for (uint32_t b = grp; b < nblocks; b += 128) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (b + 32 * i < nblocks) q[i] += partial[pslot(p, b + 32 * i, nblocks)];
    }
  }
  s_g[grp][pl] = (q[0] + q[1]) + (q[2] + q[3]);
  __syncthreads();
  if (grp != 0 || p >= (uint32_t)N_PARAMS) return;
  float t[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) t[i] = s_g[2 * i][pl] + s_g[2 * i + 1][pl];
#pragma unroll
  for (int i = 0; i < 8; ++i) t[i] = t[2 * i] + t[2 * i + 1];
#pragma unroll
  for (int i = 0; i < 4; ++i) t[i] = t[2 * i] + t[2 * i + 1];
  const float gg = (t[0] + t[1]) + (t[2] + t[3]);
  if (grad_out) {  // gradient only (no step)
    grad_out[p] = gg;
    return;
  }
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

}  // namespace psq
