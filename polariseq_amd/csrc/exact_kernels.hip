// Exact kNN for gfx950: FP32 matrix-core filter with a proven error margin,
// FP64 re-rank.
//
// Order and tie rule follow by_distance / knn_exact of the reference
// (crates/polariseq-core/src/doublets.rs:544-651): distance, then index, the
// query itself excluded. The ranking distance is FP64, sum (double(a)-double(b))^2
// summed in dimension order with no fused multiply-add, so the result is
// bit-exact against the CPU oracle that computes the same sequence.
//
// norms   xn[p] = |p|^2 in FP32 (dimension order), and P = max_p xn[p].
// pass 1  a block of four waves, each holding one or two sets of 16 queries
//         (two at the compiled widths d = 30, 50), streams a segment of points
//         through LDS, 64 at a time; per 16 points and query set one
//         v_mfma_f32_16x16x4f32 chain over the dimensions gives the 16 x 16 dot
//         products (the reference's tiled kNN also works on matrix products,
//         doublets.rs:557-651), and e(p, q) = (xn[p] + xn[q]) - 2 p.q.
//         Lane (c, g) of a wave sees query c against points 4g .. 4g + 3 of every
//         16 and keeps its K smallest e: four lists per query and segment. A
//         tile whose 16 values for the lane are all at or above the lane's bound
//         skips the insertion; the bound is the list's K-th value or, tighter,
//         the largest of the query's four lists' ceil(k/4)-th values (together
//         at least k values), which no later value at or above it can change.
// tau    the k-th smallest e over all lists -> threshold T = tau + 2 E(q), with
//         E(q) = 2.1 (d + 8) 2^-24 (xn[q] + P) + 2^-110 (xn[q] + P + 1), a bound on
//         |e(p, q) - |p - q|^2| for every p: the dot product, both norms and the
//         two additions each err by at most (d + 2) 2^-24 times |p|^2 + |q|^2 (any
//         summation order), and the last term covers denormal products flushed
//         to zero.
// pass 2  the same scan; every point with e <= T becomes a candidate.
//         The true top-k is inside: k points have e <= tau, so the true k-th
//         distance is at most tau + E, and a true neighbour p has
//         e(p) <= |p - q|^2 + E <= tau + 2 E = T.
// rerank  one wave per query ranks its candidates by (d64, index).
// fallback  a query with more candidates than the buffer (many exact ties)
//         is finished by one thread scanning every point, with the VALU FP32
//         distance (within (d + 4) 2^-24 of the exact one, relative) against T
//         scaled by 1 + 4 (d + 8) 2^-24.
#include "psq_kernels.hpp"

namespace psq {

typedef float f32x4_t __attribute__((ext_vector_type(4)));
// FP32 distance of the fallback: four FMA chains, dimensions t = 4j + r in
// chain r, the chains summed as (a0 + a1) + (a2 + a3).
__device__ inline float filt_dist_rt(const float* __restrict__ p, const float* __restrict__ q, int d) {
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
  int t = 0;
  for (; t + 4 <= d; t += 4) {
    float d0 = p[t] - q[t], d1 = p[t + 1] - q[t + 1], d2 = p[t + 2] - q[t + 2],
          d3 = p[t + 3] - q[t + 3];
    a0 = fmaf(d0, d0, a0);
    a1 = fmaf(d1, d1, a1);
    a2 = fmaf(d2, d2, a2);
    a3 = fmaf(d3, d3, a3);
  }
  if (t < d) {
    float d0 = p[t] - q[t];
    a0 = fmaf(d0, d0, a0);
  }
  if (t + 1 < d) {
    float d1 = p[t + 1] - q[t + 1];
    a1 = fmaf(d1, d1, a1);
  }
  if (t + 2 < d) {
    float d2 = p[t + 2] - q[t + 2];
    a2 = fmaf(d2, d2, a2);
  }
  return (a0 + a1) + (a2 + a3);
}

// The FP64 ranking distance, identical operation sequence to the oracle.
__device__ inline double dist64(const float* __restrict__ a, const float* __restrict__ b, int d) {
#pragma clang fp contract(off)
  double s = 0.0;
  for (int t = 0; t < d; ++t) {
    double x = (double)a[t] - (double)b[t];
    s = s + x * x;
  }
  return s;
}

// ── norms and the matrix-core filter ──
// xn[p] = sum_t x_t^2 (FP32, dimension order); pmax = max over p, as the bits of
// a non-negative float (unsigned order = value order).
__global__ void k_exact_norms(const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp,
                              float* __restrict__ xn, uint32_t* __restrict__ pmax) {
  __shared__ uint32_t s_max;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    const float* r = Xp + (size_t)p * dp;
    float a = 0.0f;
    for (uint32_t t = 0; t < d; ++t) a = fmaf(r[t], r[t], a);
    xn[p] = a;
    atomicMax(&s_max, __float_as_uint(a));
  }
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(pmax, s_max);
}

// FP64 threshold T = tau + 2 E(q), rounded up to FP32 (see the file header).
__device__ inline float exact_threshold(float tau, float xq, const uint32_t* pmax, uint32_t d) {
  if (!(tau < __builtin_inff())) return __builtin_inff();
  const double P = (double)__uint_as_float(*pmax);
  const double nrm = (double)xq + P;
  const double E = 2.1 * ((double)d + 8.0) * 5.9604644775390625e-08 * nrm +
                   7.703719777548943e-34 * (nrm + 1.0);  // 2^-110
  const double Td = (double)tau + 2.0 * E;
  float Tf = (float)Td;
  if ((double)Tf < Td) Tf = nextafterf(Tf, __builtin_inff());
  return Tf;
}


constexpr int EXM_P = EXM_P_TILE;  // points per LDS tile

// The LDS tile holds 64 points, point r at r * TS floats, its dimensions in
// order (dimensions d .. 4 KS - 1 are the zero padding of Xp). Lane group g
// takes dimensions g KS .. g KS + KS - 1 of point and query, one per matrix
// step s: word (16 t + c) TS + g KS + s of a read lands on 64 distinct banks
// when TS = 4 KS is 4 mod 8, and TS = 4 KS + 4 restores that when 4 KS is
// 0 mod 8. The 64 norms follow the tile.
__host__ __device__ constexpr uint32_t exm_ts(uint32_t ks) { return (4 * ks) % 8 == 0 ? 4 * ks + 4 : 4 * ks; }
__host__ size_t exm_smem(uint32_t d) {
  const uint32_t ks = (d + 3) / 4;
  return ((size_t)EXM_P * exm_ts(ks) + EXM_P) * sizeof(float);
}
constexpr uint32_t EXM_THREADS = 4 * WAVE;
// Waves per SIMD the fixed-width passes are compiled for (a register budget
// of 512 / w per lane); the generic width takes what it needs.
__host__ __device__ constexpr int exm_min_waves(int D, int K) { return D > 0 && K <= 32 ? 4 : 1; }
#define EXM_WAVES(D, K) __attribute__((amdgpu_waves_per_eu(exm_min_waves(D, K))))

// Rows of a tile in registers: rows [base, base + np) of Xp, four floats at a
// time, and their norms. Rows past np repeat row base + np - 1 (finite). Every
// load is issued before the first store: the addresses are clamped instead of
// guarded, and the empty asm keeps the compiler from moving a load under the
// store's condition.
template <int NV>
struct ExmRows {
  float4 v[NV];
  float nv;
};

template <int NV>
__device__ inline void exm_fetch(ExmRows<NV>& R, const float* __restrict__ Xp, const float* __restrict__ xn,
                                 uint32_t dp, uint32_t KS, uint32_t base, uint32_t np, uint32_t e0) {
  const uint32_t tot = EXM_P * KS, last = base + np - 1;
  R.nv = xn[min(base + threadIdx.x % EXM_P, last)];
#pragma unroll
  for (int h = 0; h < NV; ++h) {
    const uint32_t e = min(e0 + h * EXM_THREADS, tot - 1);
    const uint32_t r = e / KS, j = e - r * KS;
    R.v[h] = *reinterpret_cast<const float4*>(Xp + (size_t)min(base + r, last) * dp + 4 * j);
  }
}

// Store the rows into the tile; tile norms past np are +inf, so the filter
// values of the repeated rows are +inf.
template <int NV>
__device__ inline void exm_put(ExmRows<NV>& R, uint32_t KS, uint32_t TS, uint32_t np, uint32_t e0,
                               float* tile, bool norms) {
  const uint32_t tot = EXM_P * KS;
#pragma unroll
  for (int h = 0; h < NV; ++h) {
    asm volatile("" : "+v"(R.v[h].x), "+v"(R.v[h].y), "+v"(R.v[h].z), "+v"(R.v[h].w));
    const uint32_t e = e0 + h * EXM_THREADS;
    if (e < tot) {
      const uint32_t r = e / KS, j = e - r * KS;
      *reinterpret_cast<float4*>(tile + r * TS + 4 * j) = R.v[h];
    }
  }
  if (norms && threadIdx.x < (uint32_t)EXM_P) {
    asm volatile("" : "+v"(R.nv));
    tile[EXM_P * TS + threadIdx.x] = threadIdx.x < np ? R.nv : __builtin_inff();
  }
}

// Stage a whole tile, in chunks of up to four float4 per thread (one chunk at
// the fixed widths).
template <int KSM>
__device__ inline void exm_stage(const float* __restrict__ Xp, const float* __restrict__ xn, uint32_t dp,
                                 uint32_t KS, uint32_t TS, uint32_t base, uint32_t np, float* tile) {
  constexpr int NV = (EXM_P * KSM + EXM_THREADS - 1) / EXM_THREADS;
  constexpr int CH = NV < 4 ? NV : 4;
  const uint32_t tot = EXM_P * KS;
  for (uint32_t e0 = threadIdx.x; e0 < tot; e0 += CH * EXM_THREADS) {
    ExmRows<CH> R;
    exm_fetch<CH>(R, Xp, xn, dp, KS, base, np, e0);
    exm_put<CH>(R, KS, TS, np, e0, tile, e0 == threadIdx.x);
  }
}

// The lane's query registers: qf[s] = dimension g KS + s of row qi.
template <int KSM>
__device__ inline void exm_load_query(const float* __restrict__ Xp, uint32_t dp, uint32_t KS, uint64_t qi,
                                      uint32_t g, float (&qf)[KSM]) {
#pragma unroll
  for (int s = 0; s < KSM; ++s) qf[s] = (uint32_t)s < KS ? Xp[qi * dp + g * KS + s] : 0.0f;
}

// The dot products of one tile against the wave's QS query sets (lane (c, g)
// holds query c of each set): four 16-row v_mfma_f32_16x16x4f32 chains per set
// over the dimensions, interleaved; then f(q, base, np, e) per set with
// e[4 t + i] = (xn[p] + xn[q]) - 2 p.q for row p = base + 16 t + 4 g + i
// (+inf for rows past np). The clamp at 0 is left out: the error bound holds
// without it.
template <int KSM, int QS, typename F>
__device__ inline void exm_tile(const float* tile, uint32_t KS, uint32_t TS, uint32_t base, uint32_t np,
                                uint32_t c, uint32_t g, const float (&qf)[QS][KSM], const float (&xq)[QS],
                                F&& f) {
  f32x4_t acc[QS][4];
#pragma unroll
  for (int q = 0; q < QS; ++q)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[q][t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  const float* a = tile + c * TS + g * KS;
#pragma unroll
  for (int s = 0; s < KSM; ++s)
    if ((uint32_t)s < KS) {
      float av[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) av[t] = a[16 * t * TS + s];
#pragma unroll
      for (int q = 0; q < QS; ++q)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t], qf[q][s], acc[q][t], 0, 0, 0);
    }
  const float* tn = tile + EXM_P * TS;
#pragma unroll
  for (int q = 0; q < QS; ++q) {
    float e[16];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float4 nr = *reinterpret_cast<const float4*>(tn + 16 * t + 4 * g);
      e[4 * t + 0] = fmaf(-2.0f, acc[q][t][0], nr.x + xq[q]);
      e[4 * t + 1] = fmaf(-2.0f, acc[q][t][1], nr.y + xq[q]);
      e[4 * t + 2] = fmaf(-2.0f, acc[q][t][2], nr.z + xq[q]);
      e[4 * t + 3] = fmaf(-2.0f, acc[q][t][3], nr.w + xq[q]);
    }
    f(q, base, np, e);
  }
}

// Scan rows [p0, p1) in 64-row tiles: stage a tile, then its dot products and
// f. (Double-buffered staging, the next tile's loads in flight during the
// products, measured no faster: the extra registers cost the occupancy the
// two query sets need.) Every thread of the block calls this with the same
// range.
template <int D, int QS, typename F>
__device__ inline void exm_scan(const float* __restrict__ Xp, const float* __restrict__ xn, uint32_t dp,
                                uint32_t KS, uint32_t TS, uint32_t p0, uint32_t p1, float* lds, uint32_t c,
                                uint32_t g, const float (&qf)[QS][D > 0 ? (D + 3) / 4 : 64],
                                const float (&xq)[QS], F&& f) {
  constexpr int KSM = D > 0 ? (D + 3) / 4 : 64;
  for (uint32_t base = p0; base < p1; base += EXM_P) {
    const uint32_t np = min((uint32_t)EXM_P, p1 - base);
    __syncthreads();
    exm_stage<KSM>(Xp, xn, dp, KS, TS, base, np, lds);
    __syncthreads();
    exm_tile<KSM, QS>(lds, KS, TS, base, np, c, g, qf, xq, f);
  }
}

// Smallest of the lane's 16 values of a tile.
__device__ inline float exm_min16(const float (&e)[16]) {
  float m0 = fminf(fminf(e[0], e[1]), e[2]), m1 = fminf(fminf(e[3], e[4]), e[5]);
  float m2 = fminf(fminf(e[6], e[7]), e[8]), m3 = fminf(fminf(e[9], e[10]), e[11]);
  float m4 = fminf(fminf(e[12], e[13]), e[14]);
  return fminf(fminf(fminf(m0, m1), fminf(m2, m3)), fminf(m4, e[15]));
}

// Insert v into the ascending list top[0..K): entry u becomes the median of
// (top[u-1], v, top[u]).
template <int K>
__device__ inline void exm_insert(float (&top)[K], float v) {
#pragma unroll
  for (int u = K - 1; u > 0; --u) top[u] = __builtin_amdgcn_fmed3f(top[u - 1], v, top[u]);
  top[0] = fminf(top[0], v);
}

// Pass 1 of one tile for the lane's list top (K smallest values of rows
// 4 g .. 4 g + 3 of every 16, the query row qi excluded). A value at or above
// bnd cannot change the query's k smallest over its four lists: the four lanes
// of the query together hold at least k values at or below the shared bound
// (jb = ceil(k / 4) - 1 < K / 4: every list's entry jb, maximised over the four
// lanes), and the lane's own K >= k values bound its list. Only lanes with a
// value below the bound enter the insertion.
template <int K>
__device__ inline void exm_pass1_tile(float (&top)[K], float& shared_bnd, uint32_t jb, uint32_t base,
                                      uint32_t g, uint32_t qi, const float (&e)[16]) {
  const float bnd = fminf(top[K - 1], shared_bnd);
  if (exm_min16(e) < bnd) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t p = base + 16 * (j / 4) + 4 * g + (j % 4);
      if (e[j] < fminf(top[K - 1], shared_bnd) && p != qi) exm_insert<K>(top, e[j]);
    }
  }
  if (jb < (uint32_t)K / 4) {
    float b = top[0];
#pragma unroll
    for (int j = 1; j < K / 4; ++j) b = (uint32_t)j == jb ? top[j] : b;
    b = fmaxf(b, __shfl_xor(b, 16, WAVE));
    b = fmaxf(b, __shfl_xor(b, 32, WAVE));
    shared_bnd = b;
  }
}

// Pass 2 of one tile: every row with e <= thr (the query row excluded) is
// appended to the query's candidates.
__device__ inline void exm_pass2_tile(uint32_t base, uint32_t np, uint32_t g, uint32_t qi, float thr,
                                      uint32_t* cnt_q, uint32_t* cand_q, uint32_t cap, const float (&e)[16]) {
  if (exm_min16(e) <= thr) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t r = 16 * (j / 4) + 4 * g + (j % 4);
      if (e[j] <= thr && r < np && base + r != qi) {
        const uint32_t pos = atomicAdd(cnt_q, 1u);
        if (pos < cap) cand_q[pos] = base + r;
      }
    }
  }
}

template <int D, int K>
__global__ __launch_bounds__(256) EXM_WAVES(D, K) void k_exact_pass1(const float* __restrict__ Xp,
                                                     const float* __restrict__ xn, uint32_t n,
                                                     uint32_t d, uint32_t dp,
                                                     const uint32_t* __restrict__ queries,
                                                     uint32_t nq, uint32_t seg_len, uint32_t nseg,
                                                     uint32_t k, float* __restrict__ partial) {
  constexpr int KSM = D > 0 ? (D + 3) / 4 : 64;
  constexpr int QS = (int)exm_qsets(D, K);
  extern __shared__ float exm_lds[];
  const uint32_t KS = D > 0 ? (uint32_t)KSM : (d + 3) / 4, TS = exm_ts(KS);
  const uint32_t l = threadIdx.x % 64, w = threadIdx.x / 64, c = l % 16, g = l / 16;
  uint32_t qslot[QS], qi[QS];
  float qf[QS][KSM], xq[QS], top[QS][K], sb[QS];
#pragma unroll
  for (int q = 0; q < QS; ++q) {
    qslot[q] = (blockIdx.x * 4 + w) * 16 * QS + q * 16 + c;
    qi[q] = qslot[q] < nq ? (queries ? queries[qslot[q]] : qslot[q]) : 0u;
    exm_load_query<KSM>(Xp, dp, KS, qi[q], g, qf[q]);
    xq[q] = xn[qi[q]];
    sb[q] = __builtin_inff();
#pragma unroll
    for (int j = 0; j < K; ++j) top[q][j] = __builtin_inff();
  }
  const uint32_t jb = (k + 3) / 4 - 1;
  const uint32_t s0 = blockIdx.y * seg_len;
  exm_scan<D, QS>(Xp, xn, dp, KS, TS, s0, min(n, s0 + seg_len), exm_lds, c, g, qf, xq,
                  [&](int q, uint32_t base, uint32_t, const float (&e)[16]) {
                    exm_pass1_tile<K>(top[q], sb[q], jb, base, g, qi[q], e);
                  });
#pragma unroll
  for (int q = 0; q < QS; ++q)
    if (qslot[q] < nq) {
      float* out = partial + (((size_t)qslot[q] * nseg + blockIdx.y) * 4 + g) * K;
#pragma unroll
      for (int j = 0; j < K; ++j) out[j] = top[q][j];
    }
}

// tau over a query's L = 4 nseg K pass-1 values (K <= 64, k <= 64), then T.
__global__ void k_exact_tau(const float* __restrict__ partial, uint32_t nq, uint32_t L, uint32_t k,
                            uint32_t d, const uint32_t* __restrict__ queries,
                            const float* __restrict__ xn, const uint32_t* __restrict__ pmax,
                            float* __restrict__ T) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  float top[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) top[j] = __builtin_inff();
  const float* src = partial + (size_t)q * L;
  for (uint32_t e = 0; e < L; ++e) {
    float v = src[e];
    if (v < top[63]) {
#pragma unroll
      for (int t = 63; t > 0; --t) {
        float a = top[t - 1];
        top[t] = v < a ? a : (v < top[t] ? v : top[t]);
      }
      top[0] = v < top[0] ? v : top[0];
    }
  }
  float tau = __builtin_inff();
#pragma unroll
  for (int j = 0; j < 64; ++j)
    if ((uint32_t)j == k - 1) tau = top[j];
  T[q] = exact_threshold(tau, xn[queries ? queries[q] : q], pmax, d);
}

template <int D>
__global__ __launch_bounds__(256) EXM_WAVES(D, 0) void k_exact_pass2(const float* __restrict__ Xp,
                                                     const float* __restrict__ xn, uint32_t n,
                                                     uint32_t d, uint32_t dp,
                                                     const uint32_t* __restrict__ queries,
                                                     uint32_t nq, uint32_t seg_len,
                                                     const float* __restrict__ T,
                                                     uint32_t* __restrict__ cnt,
                                                     uint32_t* __restrict__ cand, uint32_t cap) {
  constexpr int KSM = D > 0 ? (D + 3) / 4 : 64;
  constexpr int QS = (int)exm_qsets(D, 0);
  extern __shared__ float exm_lds[];
  const uint32_t KS = D > 0 ? (uint32_t)KSM : (d + 3) / 4, TS = exm_ts(KS);
  const uint32_t l = threadIdx.x % 64, w = threadIdx.x / 64, c = l % 16, g = l / 16;
  uint32_t qslot[QS], qi[QS];
  float qf[QS][KSM], xq[QS], thr[QS];
#pragma unroll
  for (int q = 0; q < QS; ++q) {
    qslot[q] = (blockIdx.x * 4 + w) * 16 * QS + q * 16 + c;
    const bool ok = qslot[q] < nq;
    qi[q] = ok ? (queries ? queries[qslot[q]] : qslot[q]) : 0u;
    thr[q] = ok ? T[qslot[q]] : -1.0f;
    exm_load_query<KSM>(Xp, dp, KS, qi[q], g, qf[q]);
    xq[q] = xn[qi[q]];
  }
  const uint32_t s0 = blockIdx.y * seg_len;
  exm_scan<D, QS>(Xp, xn, dp, KS, TS, s0, min(n, s0 + seg_len), exm_lds, c, g, qf, xq,
                  [&](int q, uint32_t base, uint32_t np, const float (&e)[16]) {
                    exm_pass2_tile(base, np, g, qi[q], thr[q], &cnt[qslot[q]],
                                   cand + (size_t)qslot[q] * cap, cap, e);
                  });
}

// One wave per query: rank each candidate by (d64, index) among all of them.
__global__ __launch_bounds__(256) void k_exact_rerank(
    const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp,
    const uint32_t* __restrict__ queries, uint32_t nq, uint32_t k, const uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ cand, uint32_t cap, uint32_t* __restrict__ idx_out,
    float* __restrict__ dist_out, uint32_t* __restrict__ overflow, uint32_t* __restrict__ n_overflow) {
  constexpr int CPL = 2;  // cap <= 128
  const int w = threadIdx.x / WAVE;
  const int l = lane_id();
  const uint32_t qs = blockIdx.x * 4 + w;
  if (qs >= nq) return;  // wave-uniform
  const uint32_t c = cnt[qs];
  if (c > cap) {
    if (l == 0) overflow[atomicAdd(n_overflow, 1u)] = qs;
    return;
  }
  const uint32_t qi = queries ? queries[qs] : qs;
  const float* qrow = Xp + (size_t)qi * dp;
  double dv[CPL];
  uint32_t iv[CPL];
#pragma unroll
  for (int r = 0; r < CPL; ++r) {
    const uint32_t e = r * WAVE + l;
    if (e < c) {
      iv[r] = cand[(size_t)qs * cap + e];
      dv[r] = dist64(qrow, Xp + (size_t)iv[r] * dp, (int)d);
    } else {
      iv[r] = EMPTY;
      dv[r] = __builtin_inf();
    }
  }
  uint32_t rank[CPL] = {0, 0};
  for (uint32_t j = 0; j < c; ++j) {
    const int src = (int)(j & (WAVE - 1));
    const int rj = (int)(j / WAVE);
    double dj = 0.0;
    uint32_t ij = 0;
#pragma unroll
    for (int r = 0; r < CPL; ++r) {
      double t = __shfl(dv[r], src, WAVE);
      uint32_t u = (uint32_t)__shfl((int)iv[r], src, WAVE);
      if (r == rj) {
        dj = t;
        ij = u;
      }
    }
#pragma unroll
    for (int r = 0; r < CPL; ++r) rank[r] += (dj < dv[r] || (dj == dv[r] && ij < iv[r])) ? 1u : 0u;
  }
#pragma unroll
  for (int r = 0; r < CPL; ++r) {
    const uint32_t e = r * WAVE + l;
    if (e < c && rank[r] < k) {
      idx_out[(size_t)qs * k + rank[r]] = iv[r];
      dist_out[(size_t)qs * k + rank[r]] = (float)__dsqrt_rn(dv[r]);
    }
  }
  for (uint32_t pos = c + l; pos < k; pos += WAVE) {
    idx_out[(size_t)qs * k + pos] = EMPTY;
    dist_out[(size_t)qs * k + pos] = __builtin_inff();
  }
}

// One thread per overflowed query: scan every point, keep the k smallest
// (d64, index) among those passing the FP32 threshold (sorted insertion in a
// global scratch list).
__global__ void k_exact_fallback(const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp,
                                 const uint32_t* __restrict__ queries, uint32_t k,
                                 const float* __restrict__ T, const uint32_t* __restrict__ overflow,
                                 uint32_t n_overflow, double* __restrict__ sd,
                                 uint32_t* __restrict__ si, uint32_t* __restrict__ idx_out,
                                 float* __restrict__ dist_out, const uint64_t* __restrict__ off,
                                 uint32_t S, double thr_scale) {
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_overflow) return;
  const uint32_t qs = overflow[o];
  const uint32_t qi = queries ? queries[qs] : qs;
  const float* qrow = Xp + (size_t)qi * dp;
  // the matrix-core threshold bounds the filter value; the VALU distance here
  // is within (d + 4) 2^-24 of the exact one, so T scaled by thr_scale >= 1 + that
  // bound keeps every true neighbour (thr_scale = 1 for the batched form's own T)
  double thd = (double)T[qs] * thr_scale;
  float thr = (float)thd;
  if ((double)thr < thd) thr = nextafterf(thr, __builtin_inff());
  double* ld = sd + (size_t)o * k;
  uint32_t* li = si + (size_t)o * k;
  uint32_t cntk = 0;
  uint32_t p0 = 0, p1 = n;
  if (off) {  // batched: scan the query's own sample only
    uint32_t lo = 0, hi = S;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (off[mid] <= qi) lo = mid; else hi = mid;
    }
    p0 = (uint32_t)off[lo];
    p1 = (uint32_t)off[lo + 1];
  }
  for (uint32_t p = p0; p < p1; ++p) {
    if (p == qi) continue;
    const float* pr = Xp + (size_t)p * dp;
    if (!(filt_dist_rt(pr, qrow, (int)d) <= thr)) continue;
    double dd = dist64(qrow, pr, (int)d);
    if (cntk == k && !(dd < ld[k - 1] || (dd == ld[k - 1] && p < li[k - 1]))) continue;
    uint32_t pos = cntk < k ? cntk : k - 1;
    while (pos > 0 && (dd < ld[pos - 1] || (dd == ld[pos - 1] && p < li[pos - 1]))) {
      ld[pos] = ld[pos - 1];
      li[pos] = li[pos - 1];
      --pos;
    }
    ld[pos] = dd;
    li[pos] = p;
    if (cntk < k) ++cntk;
  }
  for (uint32_t s = 0; s < k; ++s) {
    idx_out[(size_t)qs * k + s] = s < cntk ? li[s] : EMPTY;
    dist_out[(size_t)qs * k + s] = s < cntk ? (float)__dsqrt_rn(ld[s]) : __builtin_inff();
  }
}

// ── large k (64 < k <= EX_KBIG_MAX; Scrublet's k_adj, doublets.rs:1083-1090) ──
// Pass 1 keeps each segment's 64 smallest FP32 distances as above, with enough
// segments that a query's true top k rarely crowd one segment. The k-th
// smallest of the union of the segments' lists is never below the true k-th
// smallest FP32 distance (the union is a subset of all distances), so the
// threshold built from it keeps the proof of pass 2. The selection and the
// re-rank run per query in LDS.

// tau: the k-th smallest of a query's L pass-1 values, by radix select (four
// 8-bit digits) on an order-preserving image of the FP32 bits (a value can
// round below zero: negative values have their bits inverted, the others their
// sign bit set). One block per query; dynamic LDS: L values.
__global__ __launch_bounds__(256) void k_exact_tau_big(const float* __restrict__ partial,
                                                       uint32_t nq, uint32_t L, uint32_t k, uint32_t d,
                                                       const uint32_t* __restrict__ queries,
                                                       const float* __restrict__ xn,
                                                       const uint32_t* __restrict__ pmax,
                                                       float* __restrict__ T) {
  extern __shared__ uint32_t sv[];
  __shared__ uint32_t hist[256];
  __shared__ uint32_t s_prefix, s_mask, s_rank;
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  const float* src = partial + (size_t)q * L;
  for (uint32_t e = tid; e < L; e += blockDim.x) {
    const uint32_t b = __float_as_uint(src[e]);
    sv[e] = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  }
  if (tid == 0) {
    s_prefix = 0;
    s_mask = 0;
    s_rank = k;  // 1-based rank still to find under the current prefix
  }
  __syncthreads();
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    const uint32_t prefix = s_prefix, mask = s_mask;
    for (uint32_t e = tid; e < L; e += blockDim.x) {
      const uint32_t v = sv[e];
      if ((v & mask) == prefix) atomicAdd(&hist[(v >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t r = s_rank, b = 0;
      for (; b < 255u && hist[b] < r; ++b) r -= hist[b];
      s_rank = r;
      s_prefix = prefix | (b << shift);
      s_mask = mask | (255u << shift);
    }
    __syncthreads();
  }
  if (tid != 0) return;
  const uint32_t u = s_prefix;
  const float tau = __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
  T[q] = exact_threshold(tau, xn[queries ? queries[q] : q], pmax, d);
}

// One block per query: the FP64 distance of every candidate, a bitonic sort of
// (d64 bits, index) in LDS, the first k written. Dynamic LDS: P u64 + P u32,
// P = cap rounded up to a power of two.
__global__ __launch_bounds__(256) void k_exact_rerank_big(
    const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp,
    const uint32_t* __restrict__ queries, uint32_t nq, uint32_t k, const uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ cand, uint32_t cap, uint32_t pmax, uint32_t* __restrict__ idx_out,
    float* __restrict__ dist_out, uint32_t* __restrict__ overflow,
    uint32_t* __restrict__ n_overflow) {
  extern __shared__ uint64_t s_d[];
  uint32_t* s_i = reinterpret_cast<uint32_t*>(s_d + pmax);
  const uint32_t qs = blockIdx.x, tid = threadIdx.x;
  const uint32_t c = cnt[qs];
  if (c > cap) {
    if (tid == 0) overflow[atomicAdd(n_overflow, 1u)] = qs;
    return;
  }
  const uint32_t qi = queries ? queries[qs] : qs;
  const float* qrow = Xp + (size_t)qi * dp;
  uint32_t P = 1;
  while (P < c) P <<= 1;
  for (uint32_t e = tid; e < P; e += blockDim.x) {
    if (e < c) {
      const uint32_t i = cand[(size_t)qs * cap + e];
      s_d[e] = (uint64_t)__double_as_longlong(dist64(qrow, Xp + (size_t)i * dp, (int)d));
      s_i[e] = i;
    } else {
      s_d[e] = ~0ull;
      s_i[e] = EMPTY;
    }
  }
  __syncthreads();
  for (uint32_t size = 2; size <= P; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = tid; t < P / 2; t += blockDim.x) {
        const uint32_t lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t da = s_d[lo], db = s_d[hi];
        const uint32_t ia = s_i[lo], ib = s_i[hi];
        const bool gt = da > db || (da == db && ia > ib);
        if (gt == up) {
          s_d[lo] = db;
          s_d[hi] = da;
          s_i[lo] = ib;
          s_i[hi] = ia;
        }
      }
      __syncthreads();
    }
  for (uint32_t pos = tid; pos < k; pos += blockDim.x) {
    const bool has = pos < c;
    idx_out[(size_t)qs * k + pos] = has ? s_i[pos] : EMPTY;
    dist_out[(size_t)qs * k + pos] =
        has ? (float)__dsqrt_rn(__longlong_as_double((long long)s_d[pos])) : __builtin_inff();
  }
}

// q[i] = q0 + i
__global__ void k_iota_u32(uint32_t* __restrict__ q, uint32_t q0, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) q[i] = q0 + i;
}

template __global__ void k_exact_pass1<0, 16>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<0, 32>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<0, 64>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<30, 16>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<30, 32>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<30, 64>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<50, 16>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<50, 32>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass1<50, 64>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, float*);
template __global__ void k_exact_pass2<0>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, const float*, uint32_t*, uint32_t*, uint32_t);
template __global__ void k_exact_pass2<30>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, const float*, uint32_t*, uint32_t*, uint32_t);
template __global__ void k_exact_pass2<50>(const float*, const float*, uint32_t, uint32_t, uint32_t,
    const uint32_t*, uint32_t, uint32_t, const float*, uint32_t*, uint32_t*, uint32_t);

}  // namespace psq

// ── batched form: many independent samples in one persistent launch ──
//
// Samples are contiguous row ranges [off[s], off[s+1]) of one array. Every
// point of a sample is a query against the points of its own sample (self
// excluded), as Scrublet's per-sample kNN does (crates/polariseq-core/src/
// doublets.rs:1085-1113, 1366-1374). Work units (sample, query block, point
// segment) are pulled from an atomic counter by a grid sized to the device;
// every block leaves the loop once the counter passes the unit count. A unit's
// 64 queries run the matrix-core filter of the single-set form above; the
// segment length comes from the autotuner (batched.hip).
namespace psq {

template <int D, int K>
__global__ __launch_bounds__(256) EXM_WAVES(D, K) void k_bexact_pass1(
    const float* __restrict__ Xp, const float* __restrict__ xn, uint32_t d, uint32_t dp,
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ units, uint32_t U, uint32_t seg_len,
    const uint32_t* __restrict__ nseg_of, const uint64_t* __restrict__ part_off, uint32_t k,
    uint32_t* __restrict__ counter, float* __restrict__ partial) {
  constexpr int KSM = D > 0 ? (D + 3) / 4 : 64;
  constexpr int QS = (int)exm_qsets(D, K);
  extern __shared__ float exm_lds[];
  __shared__ uint32_t s_u;
  const uint32_t KS = D > 0 ? (uint32_t)KSM : (d + 3) / 4, TS = exm_ts(KS);
  const uint32_t l = threadIdx.x % 64, w = threadIdx.x / 64, c = l % 16, g = l / 16;
  const uint32_t jb = (k + 3) / 4 - 1;
  for (;;) {
    if (threadIdx.x == 0) s_u = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t u = s_u;
    __syncthreads();
    if (u >= U) break;
    const uint32_t s = units[3 * u], qb = units[3 * u + 1], seg = units[3 * u + 2];
    const uint32_t o = (uint32_t)off[s];
    const uint32_t ns = (uint32_t)off[s + 1] - o;
    uint32_t ql[QS], qi[QS];
    float qf[QS][KSM], xq[QS], top[QS][K], sb[QS];
#pragma unroll
    for (int q = 0; q < QS; ++q) {
      ql[q] = (qb * 4 + w) * 16 * QS + q * 16 + c;
      qi[q] = o + (ql[q] < ns ? ql[q] : 0u);
      exm_load_query<KSM>(Xp, dp, KS, qi[q], g, qf[q]);
      xq[q] = xn[qi[q]];
      sb[q] = __builtin_inff();
#pragma unroll
      for (int j = 0; j < K; ++j) top[q][j] = __builtin_inff();
    }
    const uint32_t s0 = o + seg * seg_len;
    exm_scan<D, QS>(Xp, xn, dp, KS, TS, s0, o + min(ns, seg * seg_len + seg_len), exm_lds, c, g, qf, xq,
                    [&](int q, uint32_t base, uint32_t, const float (&e)[16]) {
                      exm_pass1_tile<K>(top[q], sb[q], jb, base, g, qi[q], e);
                    });
#pragma unroll
    for (int q = 0; q < QS; ++q)
      if (ql[q] < ns) {
        float* out = partial + part_off[s] + (((size_t)ql[q] * nseg_of[s] + seg) * 4 + g) * K;
#pragma unroll
        for (int j = 0; j < K; ++j) out[j] = top[q][j];
      }
  }
}

// tau per global query: the k-th smallest over its sample's 4 nseg lists, then T.
__global__ void k_bexact_tau(const float* __restrict__ partial, const uint64_t* __restrict__ off,
                             uint32_t S, const uint32_t* __restrict__ nseg_of,
                             const uint64_t* __restrict__ part_off, uint64_t nq, uint32_t K,
                             uint32_t k, uint32_t d, const float* __restrict__ xn,
                             const uint32_t* __restrict__ pmax, float* __restrict__ T) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nq) return;
  uint32_t lo = 0, hi = S;  // sample of g: off[lo] <= g < off[lo + 1]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (off[mid] <= g) lo = mid; else hi = mid;
  }
  const uint32_t s = lo;
  const uint64_t ql = g - off[s];
  const uint32_t L = 4 * nseg_of[s] * K;
  float top[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) top[j] = __builtin_inff();
  const float* src = partial + part_off[s] + ql * L;
  for (uint32_t e = 0; e < L; ++e) {
    const float v = src[e];
    if (v < top[63]) {
#pragma unroll
      for (int t = 63; t > 0; --t) {
        const float a = top[t - 1];
        top[t] = v < a ? a : (v < top[t] ? v : top[t]);
      }
      top[0] = v < top[0] ? v : top[0];
    }
  }
  float tau = __builtin_inff();
#pragma unroll
  for (int j = 0; j < 64; ++j)
    if ((uint32_t)j == k - 1) tau = top[j];
  T[g] = exact_threshold(tau, xn[g], pmax, d);
}

template <int D>
__global__ __launch_bounds__(256) EXM_WAVES(D, 0) void k_bexact_pass2(
    const float* __restrict__ Xp, const float* __restrict__ xn, uint32_t d, uint32_t dp,
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ units, uint32_t U, uint32_t seg_len,
    const float* __restrict__ T, uint32_t* __restrict__ counter, uint32_t* __restrict__ cnt,
    uint32_t* __restrict__ cand, uint32_t cap) {
  constexpr int KSM = D > 0 ? (D + 3) / 4 : 64;
  constexpr int QS = (int)exm_qsets(D, 0);
  extern __shared__ float exm_lds[];
  __shared__ uint32_t s_u;
  const uint32_t KS = D > 0 ? (uint32_t)KSM : (d + 3) / 4, TS = exm_ts(KS);
  const uint32_t l = threadIdx.x % 64, w = threadIdx.x / 64, c = l % 16, g = l / 16;
  for (;;) {
    if (threadIdx.x == 0) s_u = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t u = s_u;
    __syncthreads();
    if (u >= U) break;
    const uint32_t s = units[3 * u], qb = units[3 * u + 1], seg = units[3 * u + 2];
    const uint32_t o = (uint32_t)off[s];
    const uint32_t ns = (uint32_t)off[s + 1] - o;
    uint32_t qi[QS];
    float qf[QS][KSM], xq[QS], thr[QS];
#pragma unroll
    for (int q = 0; q < QS; ++q) {
      const uint32_t ql = (qb * 4 + w) * 16 * QS + q * 16 + c;
      qi[q] = o + (ql < ns ? ql : 0u);
      thr[q] = ql < ns ? T[qi[q]] : -1.0f;
      exm_load_query<KSM>(Xp, dp, KS, qi[q], g, qf[q]);
      xq[q] = xn[qi[q]];
    }
    const uint32_t s0 = o + seg * seg_len;
    exm_scan<D, QS>(Xp, xn, dp, KS, TS, s0, o + min(ns, seg * seg_len + seg_len), exm_lds, c, g, qf, xq,
                    [&](int q, uint32_t base, uint32_t np, const float (&e)[16]) {
                      exm_pass2_tile(base, np, g, qi[q], thr[q], &cnt[qi[q]], cand + (size_t)qi[q] * cap,
                                     cap, e);
                    });
  }
}

template __global__ void k_bexact_pass1<0, 16>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<0, 32>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<0, 64>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<30, 16>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<30, 32>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<30, 64>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<50, 16>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<50, 32>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass1<50, 64>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint64_t*, uint32_t, uint32_t*, float*);
template __global__ void k_bexact_pass2<0>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const float*, uint32_t*, uint32_t*, uint32_t*, uint32_t);
template __global__ void k_bexact_pass2<30>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const float*, uint32_t*, uint32_t*, uint32_t*, uint32_t);
template __global__ void k_bexact_pass2<50>(const float*, const float*, uint32_t, uint32_t,
    const uint64_t*, const uint32_t*, uint32_t, uint32_t, const float*, uint32_t*, uint32_t*, uint32_t*, uint32_t);

}  // namespace psq
