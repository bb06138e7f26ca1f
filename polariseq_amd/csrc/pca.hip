// PCA of a CSR matrix by the exact Gram method, on the GPU
// (crates/polariseq-core/src/pca.rs:396-545, rsvd.rs:155-177, 241-283, 356-390, 439-456):
//
//   optional scaled view (pp.scale semantics, pca.rs:396-480): stored value x of
//     gene j becomes clip((x - mu_j) / sd_j) - b_j, b_j = clip(-mu_j / sd_j), the
//     gene statistics in f64 with ddof 1 (preprocess.rs:314-358); scaling centres;
//   C = S^T S in f64, then C_ij -= (mu_i n) mu_j with mu the f32 column means and
//     i >= j (the triangle the reference's solver reads), mirrored;
//   the top k eigenpairs of C (f64); V = eigenvectors as f32 (row-major m x k);
//   embedding = S V - 1 (mu^T V), f32, each row summed over its stored entries in
//     order with separate multiply and add (rsvd.rs:241-283); every column
//     flipped so that its first largest-magnitude entry is positive
//     (rsvd.rs:439-456);
//   variance ratio = max(lambda_j, 0) / trace(C).
//
// The reference takes the eigendecomposition with a dense solver; here the top
// k come from Chebyshev-accelerated subspace iteration (Rutishauser's
// simultaneous iteration, Numer. Math. 16 (1970) 205-223) on A = C + sigma I
// (sigma = 1e-4 trace(C) / m keeps every block full rank; the eigenvectors are
// C's), block b = k + max(16, k / 2): a Rayleigh-Ritz step (cyclic Jacobi in
// round-robin order on the b x b projection, one workgroup), then a Chebyshev
// filter T_d(A / e) on [-e, e], e the block's smallest Ritz value, with the
// converged leading Ritz pairs locked (left out of the filter and deflated from
// A), then shifted CholeskyQR3 (Fukaya, Kannan, Nakatsukasa, Yamamoto,
// Yanagisawa, SIAM J. Sci. Comput. 42 (2020) A477), until every wanted residual
// is below 1e-11 of the largest eigenvalue. The Gram of densified row chunks,
// the products with A and the small products are rocBLAS f64 calls with
// atomics disabled; every other reduction has a fixed order, so a run repeats
// bit for bit.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/psq_knn.h"
#include "psq_kernels.hpp"

namespace psq_host {
int fail(int code, const char* fmt, ...);
}
using psq_host::fail;

#define PHIPC(call)                                                                                  \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      return fail(e_ == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE, "%s failed: %s", #call, \
                  hipGetErrorString(e_));                                                            \
  } while (0)
#define PBLAS(call)                                                                     \
  do {                                                                                  \
    rocblas_status s_ = (call);                                                         \
    if (s_ != rocblas_status_success)                                                   \
      return fail(PSQ_E_DEVICE, "%s failed: %s", #call, rocblas_status_to_string(s_)); \
  } while (0)
#define PTRY(expr)                 \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != PSQ_OK) return rc_; \
  } while (0)

namespace psq {

// The reference takes the Gram route up to 4096 genes (pca.rs:201) and randomized
// SVD above; this build stays exact up to 32768 (C and the padded Gram: 8 GiB of
// f64 each in HBM; m^2 stays below 2^32 work-items for the one-thread-per-entry
// passes).
constexpr uint32_t PCA_MAX_FEATURES = 32768;
constexpr uint32_t PCA_MAX_BLOCK = 1024;          // subspace block b
constexpr uint32_t PCA_LDS_BLOCK = 80;            // b x b f64 (50 KB) kept in LDS
constexpr uint32_t PCA_SEGMENTS = 64;             // row segments of the column sums
constexpr uint32_t PCA_MAX_PRODUCTS = 20000;      // block products with C before giving up
constexpr int PCA_MAX_DEGREE = 24;                // Chebyshev filter degree cap

// ── input check ──

// bad |= 1 for a column index >= m, 2 for a decreasing indptr.
__global__ void k_pca_check(const uint64_t* __restrict__ indptr, uint64_t n,
                            const uint32_t* __restrict__ indices, uint64_t nnz, uint32_t m,
                            uint32_t* __restrict__ bad) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnz || t < n; t += step) {
    if (t < nnz && indices[t] >= m) atomicOr(bad, 1u);
    if (t < n && indptr[t + 1] < indptr[t]) atomicOr(bad, 2u);
  }
}

// ── column sums and the scaled view ──

// Rows [r0, r0 + R) as dense f64 rows of ld >= m + 1 elements: the stored
// values at their columns, 1 at column m (its Gram row holds the column sums),
// the chunk zero on entry. One wave per row, lanes over its entries; a column
// stored twice in a row (not canonical CSR) gets both values added.
__global__ void k_pca_densify(const uint64_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
                              const float* __restrict__ val, uint64_t r0, uint32_t R, uint32_t m, uint32_t ld,
                              double* __restrict__ D) {
  const uint32_t r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (r >= R) return;
  double* row = D + (size_t)r * ld;
  if (lane == 0) row[m] = 1.0;
  for (uint64_t e = indptr[r0 + r] + lane; e < indptr[r0 + r + 1]; e += 64)
    unsafeAtomicAdd(&row[indices[e]], (double)val[e]);
}

// Column sums and sums of squares of a dense chunk (rows of ld), per row
// segment: part[p][j] over rows [p seg, (p + 1) seg). Grid (ceil(m / 256), P).
__global__ void k_pca_colsums_part(const double* __restrict__ D, uint32_t R, uint32_t m, uint32_t ld, uint32_t seg,
                                   double* __restrict__ ps_s, double* __restrict__ ps_q) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t p = blockIdx.y;
  if (j >= m) return;
  const uint32_t lo = p * seg, hi = lo + seg < R ? lo + seg : R;
  double a = 0.0, b = 0.0;
  for (uint32_t r = lo; r < hi; ++r) {
    const double x = D[(size_t)r * ld + j];
    a += x;
    b += x * x;
  }
  ps_s[(size_t)p * m + j] = a;
  ps_q[(size_t)p * m + j] = b;
}
// acc[j] += the P partials in segment order.
__global__ void k_pca_colsums_fold(const double* __restrict__ ps_s, const double* __restrict__ ps_q,
                                   uint32_t P, uint32_t m, double* __restrict__ s, double* __restrict__ q) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double a = 0.0, b = 0.0;
  for (uint32_t p = 0; p < P; ++p) {
    a += ps_s[(size_t)p * m + j];
    b += ps_q[(size_t)p * m + j];
  }
  s[j] += a;
  q[j] += b;
}

// The m x m Gram matrix and the column sums out of the (ld x ld) product of
// the chunks with their ones column (only its upper triangle is computed, and
// only C's upper triangle is read later): C[i][j] = G[i][j], s[j] = G[j][m].
__global__ void k_pca_gram_unpack(const double* __restrict__ G, uint32_t m, uint32_t ld, double* __restrict__ C,
                                  double* __restrict__ s) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)m * m) return;
  const uint32_t col = (uint32_t)(t / m), row = (uint32_t)(t % m);
  C[t] = G[(size_t)col * ld + row];
  if (row == 0) s[col] = G[(size_t)m * ld + col];  // (col, m): upper triangle
}

// Gene statistics -> the scaled view's mean, 1/sd and clipped zero
// (pca.rs:403-414; preprocess.rs:344-358: var = (q/n - mean^2) n/(n-1), >= 0).
__global__ void k_pca_scale_stats(const double* __restrict__ s, const double* __restrict__ q, uint32_t m,
                                  uint64_t n, double clip, double* __restrict__ mean,
                                  double* __restrict__ inv_sd, double* __restrict__ zero) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const double nf = (double)n;
  const double mu = s[j] / nf;
  const double var = n < 2 ? 0.0 : fmax((q[j] / nf - mu * mu) * nf / (nf - 1.0), 0.0);
  const double sd = sqrt(var);
  const double inv = sd == 0.0 ? 1.0 : 1.0 / sd;
  mean[j] = mu;
  inv_sd[j] = inv;
  zero[j] = fmin(fmax(-mu * inv, -clip), clip);
}

// Stored values of the scaled view (pca.rs:450-460).
__global__ void k_pca_scale_values(const uint32_t* __restrict__ indices, const float* __restrict__ x,
                                   uint64_t nnz, const double* __restrict__ mean,
                                   const double* __restrict__ inv_sd, const double* __restrict__ zero,
                                   double clip, float* __restrict__ out) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += step) {
    const uint32_t j = indices[e];
    const double z = fmin(fmax(((double)x[e] - mean[j]) * inv_sd[j], -clip), clip);
    out[e] = (float)(z - zero[j]);
  }
}

// ── Gram matrix ──

__global__ void k_pca_means(const double* __restrict__ s, uint32_t m, uint64_t n, float* __restrict__ mu) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) mu[j] = (float)(s[j] / (double)n);
}

// C (column-major; the upper triangle of the Gram product) -> full centred matrix. Element
// (row, col), row <= col, is the reference's lower-triangle element (col, row):
// C - (mu_col n) mu_row.
__global__ void k_pca_center(double* __restrict__ C, uint32_t m, uint64_t n, const float* __restrict__ mu) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)m * m) return;
  const uint32_t col = (uint32_t)(t / m), row = (uint32_t)(t - (uint64_t)col * m);
  if (row > col) return;
  double v = C[t];
  if (mu) v -= ((double)mu[col] * (double)n) * (double)mu[row];
  C[t] = v;
  C[(size_t)row * m + col] = v;
}

// trace(C) in index order (pca.rs:510) and sigma = 1e-4 trace / m.
__global__ void k_pca_trace(const double* __restrict__ C, uint32_t m, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double t = 0.0;
  for (uint32_t i = 0; i < m; ++i) t += C[(size_t)i * m + i];
  out[0] = t;
  out[1] = 1e-4 * t / (double)m;
}

// ── eigensolver ──

// Start block Q (column-major m x b), uniform in [-1, 1) from a hash of the
// element index.
__global__ void k_pca_start(double* __restrict__ Q, uint32_t m, uint32_t b) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)m * b) return;
  Q[t] = (double)(splitmix(0x9E3779B97F4A7C15ull ^ splitmix(t + 0x5EEDull)) >> 11) * 0x1.0p-52 - 1.0;
}

// Y = a Y + b X, elementwise (X may be Y).
__global__ void k_pca_axpby(double* Y, double a, const double* X, double b, uint64_t count) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) Y[t] = a * Y[t] + b * X[t];
}

// H[i][j] *= lam[i] (column-major nl x nb, count = nl nb).
__global__ void k_pca_scale_rows(double* __restrict__ H, const double* __restrict__ lam, uint32_t nl,
                                 uint64_t count) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) H[t] *= lam[t % nl];
}

// The strict lower triangle of the b x b matrix T (column-major) from its upper one.
__global__ void k_pca_mirror(double* __restrict__ T, uint32_t b) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)b * b) return;
  const uint32_t col = (uint32_t)(t / b), row = (uint32_t)(t % b);
  if (row > col) T[t] = T[(size_t)row * b + col];
}

// Columns c of Q (m x b, column-major) with dead[c] set to zero. Grid (x, b).
__global__ void k_pca_zero_dead(double* __restrict__ Q, uint32_t m, const uint32_t* __restrict__ dead) {
  const uint32_t c = blockIdx.y;
  if (!dead[c]) return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    Q[(size_t)c * m + i] = 0.0;
}

// Z += sigma Q (the shifted operator), elementwise.
__global__ void k_pca_shift(double* __restrict__ Z, const double* __restrict__ Q, uint64_t count,
                            const double* __restrict__ tv) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) Z[t] += tv[1] * Q[t];
}

// One workgroup: upper Cholesky factor of the b x b matrix G (column-major,
// upper triangle read) plus shift I, in place; the strict lower triangle is
// zeroed. shifted: shift = 11 (m b + b (b + 1)) u trace(G) (the shifted
// CholeskyQR bound, u = 2^-53). Pivots are floored at 1e-28 of the largest
// diagonal (never reached for a full-rank block). Unshifted, a pivot at or
// below 16 b u of the largest diagonal marks its column dependent on the
// earlier ones (below what the Gram resolves): dead[k] = 1, row k of R
// becomes the unit row, so nothing is divided by it, and the caller zeroes
// the column; the others get dead[k] = 0. With b <= PCA_LDS_BLOCK the
// factorisation runs in LDS.
__global__ __launch_bounds__(256) void k_pca_chol(double* __restrict__ G, uint32_t b, uint32_t m,
                                                  int shifted, uint32_t* __restrict__ dead) {
  extern __shared__ double lds[];
  __shared__ double s_piv, s_shift, s_floor, s_tol;
  __shared__ int s_dead;
  const uint32_t tid = threadIdx.x;
  const bool in_lds = b <= PCA_LDS_BLOCK;
  double* A = in_lds ? lds : G;
  if (in_lds)
    for (uint32_t t = tid; t < b * b; t += blockDim.x) A[t] = G[t];
  __syncthreads();
  if (tid == 0) {
    double tr = 0.0, mx = 0.0;
    for (uint32_t i = 0; i < b; ++i) {
      tr += A[(size_t)i * b + i];
      mx = fmax(mx, A[(size_t)i * b + i]);
    }
    s_shift = shifted ? 11.0 * ((double)m * b + (double)b * (b + 1)) * 0x1.0p-53 * tr : 0.0;
    s_floor = fmax(mx * 1e-28, 1e-290);
    s_tol = shifted ? -1.0 : 16.0 * (double)b * 0x1.0p-53 * mx;
  }
  __syncthreads();
  const double shift = s_shift, floor_ = s_floor, tol = s_tol;
  for (uint32_t k = 0; k < b; ++k) {
    if (tid == 0) {
      double d = A[(size_t)k * b + k] + shift;
      const int dk = d <= tol;
      d = d > floor_ ? d : floor_;
      const double r = dk ? 1.0 : sqrt(d);
      A[(size_t)k * b + k] = r;
      s_piv = r;
      s_dead = dk;
      dead[k] = (uint32_t)dk;
    }
    __syncthreads();
    if (s_dead) {  // uniform: the unit row, no update of the rest
      for (uint32_t j = k + 1 + tid; j < b; j += blockDim.x) A[(size_t)j * b + k] = 0.0;
      __syncthreads();
      continue;
    }
    const double r = s_piv;
    for (uint32_t j = k + 1 + tid; j < b; j += blockDim.x) A[(size_t)j * b + k] /= r;  // R[k][j]
    __syncthreads();
    const uint32_t rem = b - k - 1;
    for (uint32_t t = tid; t < rem * rem; t += blockDim.x) {
      const uint32_t i = k + 1 + t / rem, j = k + 1 + t % rem;
      if (i <= j) A[(size_t)j * b + i] -= A[(size_t)i * b + k] * A[(size_t)j * b + k];
    }
    __syncthreads();
  }
  for (uint32_t t = tid; t < b * b; t += blockDim.x) {
    const uint32_t col = t / b, row = t % b;
    G[t] = row > col ? 0.0 : A[t];
  }
}

// One workgroup: eigendecomposition of the symmetric b x b matrix T
// (column-major; symmetrised first as (T + T^T) / 2) by cyclic Jacobi in
// round-robin order (b/2 disjoint rotations per round, rows, then columns and
// the accumulated rotations W), until the off-diagonal mass is at most 1e-30 of
// the total or 50 sweeps. The eigenvalues, sorted descending (ties by index),
// go to lam, the matching columns of W to Wout. T and W are overwritten. With
// b <= PCA_LDS_BLOCK, T lives in LDS.
__global__ __launch_bounds__(256) void k_pca_jacobi(double* __restrict__ Tg, uint32_t b,
                                                    double* __restrict__ W, double* __restrict__ Wout,
                                                    double* __restrict__ lam, double* __restrict__ cs) {
  extern __shared__ double lds[];
  __shared__ double s_part[256];
  __shared__ int s_done;
  __shared__ uint32_t s_order[PCA_MAX_BLOCK];
  const uint32_t tid = threadIdx.x;
  const bool in_lds = b <= PCA_LDS_BLOCK;
  double* T = in_lds ? lds : Tg;
  for (uint32_t t = tid; t < b * b; t += blockDim.x) {
    const uint32_t col = t / b, row = t % b;
    if (in_lds) T[t] = 0.5 * (Tg[t] + Tg[(size_t)row * b + col]);
    W[t] = row == col ? 1.0 : 0.0;
  }
  __syncthreads();
  if (!in_lds) {
    // symmetrise in place: the thread of (row, col), row < col, owns both halves
    for (uint32_t t = tid; t < b * b; t += blockDim.x) {
      const uint32_t col = t / b, row = t % b;
      if (row < col) {
        const double v = 0.5 * (T[t] + T[(size_t)row * b + col]);
        T[t] = v;
        T[(size_t)row * b + col] = v;
      }
    }
    __syncthreads();
  }
  const uint32_t bb = b + (b & 1);  // an even number of players; index b is a dummy
  const uint32_t half = bb / 2;
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (uint32_t t = tid; t < b * b; t += blockDim.x) {
      const double v = T[t];
      tot += v * v;
      if (t / b != t % b) off += v * v;
    }
    s_part[tid] = off;
    __syncthreads();
    if (tid == 0) {
      double a = 0.0;
      for (uint32_t i = 0; i < blockDim.x; ++i) a += s_part[i];
      s_part[0] = a;
    }
    __syncthreads();
    off = s_part[0];
    __syncthreads();
    s_part[tid] = tot;
    __syncthreads();
    if (tid == 0) {
      double a = 0.0;
      for (uint32_t i = 0; i < blockDim.x; ++i) a += s_part[i];
      s_done = !(off > 1e-30 * a);
    }
    __syncthreads();
    if (s_done) break;
    for (uint32_t rnd = 0; rnd + 1 < bb; ++rnd) {
      for (uint32_t i = tid; i < half; i += blockDim.x) {
        uint32_t p = i == 0 ? 0u : 1u + (i - 1 + rnd) % (bb - 1);
        uint32_t q = 1u + (bb - 2 - i + rnd) % (bb - 1);
        if (p > q) {
          const uint32_t x = p;
          p = q;
          q = x;
        }
        double c = 1.0, s = 0.0;
        if (q < b) {
          const double apq = T[(size_t)q * b + p];
          if (apq != 0.0) {
            const double app = T[(size_t)p * b + p], aqq = T[(size_t)q * b + q];
            const double tau = (aqq - app) / (2.0 * apq);
            const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tt * tt);
            s = tt * c;
          }
        }
        cs[4 * i] = (double)p;
        cs[4 * i + 1] = (double)q;
        cs[4 * i + 2] = c;
        cs[4 * i + 3] = s;
      }
      __syncthreads();
      for (uint32_t t = tid; t < half * b; t += blockDim.x) {  // rows: T <- J^T T
        const uint32_t i = t / b, k = t % b;
        const uint32_t p = (uint32_t)cs[4 * i], q = (uint32_t)cs[4 * i + 1];
        if (q >= b) continue;
        const double c = cs[4 * i + 2], s = cs[4 * i + 3];
        const double tp = T[(size_t)k * b + p], tq = T[(size_t)k * b + q];
        T[(size_t)k * b + p] = c * tp - s * tq;
        T[(size_t)k * b + q] = s * tp + c * tq;
      }
      __syncthreads();
      for (uint32_t t = tid; t < half * b; t += blockDim.x) {  // columns: T <- T J, W <- W J
        const uint32_t i = t / b, k = t % b;
        const uint32_t p = (uint32_t)cs[4 * i], q = (uint32_t)cs[4 * i + 1];
        if (q >= b) continue;
        const double c = cs[4 * i + 2], s = cs[4 * i + 3];
        const double tp = T[(size_t)p * b + k], tq = T[(size_t)q * b + k];
        T[(size_t)p * b + k] = c * tp - s * tq;
        T[(size_t)q * b + k] = s * tp + c * tq;
        const double wp = W[(size_t)p * b + k], wq = W[(size_t)q * b + k];
        W[(size_t)p * b + k] = c * wp - s * wq;
        W[(size_t)q * b + k] = s * wp + c * wq;
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    for (uint32_t i = 0; i < b; ++i) s_order[i] = i;
    for (uint32_t i = 1; i < b; ++i) {  // insertion sort, descending, ties by index
      const uint32_t x = s_order[i];
      const double v = T[(size_t)x * b + x];
      uint32_t j = i;
      while (j > 0) {
        const uint32_t y = s_order[j - 1];
        const double u = T[(size_t)y * b + y];
        if (u > v || (u == v && y < x)) break;
        s_order[j] = y;
        --j;
      }
      s_order[j] = x;
    }
    for (uint32_t i = 0; i < b; ++i) lam[i] = T[(size_t)s_order[i] * b + s_order[i]];
  }
  __syncthreads();
  for (uint32_t t = tid; t < b * b; t += blockDim.x) Wout[t] = W[(size_t)s_order[t / b] * b + t % b];
}

// Residual norms |Z_c - lam_c Q_c| of the first k Ritz pairs (column-major m x b),
// one block per pair, fixed-order tree sum.
__global__ __launch_bounds__(256) void k_pca_residuals(const double* __restrict__ Z,
                                                       const double* __restrict__ Q,
                                                       const double* __restrict__ lam, uint32_t m,
                                                       double* __restrict__ res) {
  const uint32_t c = blockIdx.x;
  __shared__ double part[256];
  double a = 0.0;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
    const double r = Z[(size_t)c * m + i] - lam[c] * Q[(size_t)c * m + i];
    a += r * r;
  }
  part[threadIdx.x] = a;
  __syncthreads();
  for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) res[c] = sqrt(part[0]);
}

// The k unit vectors e_{m-1-j} with eigenvalue 0: what a dense solver returns
// for the zero matrix (ascending order, reversed).
__global__ void k_pca_unit_block(double* __restrict__ Q, double* __restrict__ lam, uint32_t m, uint32_t k) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)m * k) return;
  const uint32_t c = (uint32_t)(t / m), i = (uint32_t)(t % m);
  Q[t] = i == m - 1 - c ? 1.0 : 0.0;
  if (i == 0) lam[c] = 0.0;
}

// ── projection ──

// V (row-major m x k f32) from the first k columns of Q (column-major f64).
__global__ void k_pca_vecs(const double* __restrict__ Q, uint32_t m, uint32_t k, float* __restrict__ V) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)m * k) return;
  const uint32_t j = (uint32_t)(t / k), c = (uint32_t)(t % k);
  V[t] = (float)Q[(size_t)c * m + j];
}
// mub[c] = sum_j mu_j V[j][c] in f64, j in order, as f32 (rsvd.rs:252-262).
__global__ void k_pca_mub(const float* __restrict__ mu, const float* __restrict__ V, uint32_t m, uint32_t k,
                          float* __restrict__ mub) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  double a = 0.0;
  for (uint32_t j = 0; j < m; ++j) a += (double)mu[j] * (double)V[(size_t)j * k + c];
  mub[c] = (float)a;
}

// out[i][c] = (sum over row i's stored entries e, in order, of s_e V[j_e][c])
// - mub[c], f32 with separate multiply and add (rsvd.rs:265-280). One thread per
// (row, component).
__global__ void k_pca_project(const uint64_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
                              const float* __restrict__ s, uint64_t n, const float* __restrict__ V,
                              uint32_t k, const float* __restrict__ mub, float* __restrict__ out) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * k; t += step) {
    const uint64_t i = t / k;
    const uint32_t c = (uint32_t)(t - i * k);
    float acc = 0.0f;
    for (uint64_t e = indptr[i]; e < indptr[i + 1]; ++e) acc = acc + s[e] * V[(size_t)indices[e] * k + c];
    if (mub) acc = acc - mub[c];
    out[t] = acc;
  }
}

// Column signs (rsvd.rs:439-456): the first entry of largest magnitude (over
// magnitudes > 0; row 0 if none) becomes positive.
// Pass 1: block g scans rows [g seg, (g + 1) seg); thread t takes column t % k
// and rows t / k + rps i. Partial (largest |u|, first row) per (block, column).
__global__ void k_pca_sign_part(const float* __restrict__ U, uint64_t n, uint32_t k, uint32_t rps,
                                uint64_t seg, float* __restrict__ pv, uint64_t* __restrict__ pi) {
  __shared__ float s_v[1024];
  __shared__ uint64_t s_i[1024];
  const uint32_t tid = threadIdx.x, c = tid % k, r0 = tid / k;
  const uint64_t lo = (uint64_t)blockIdx.x * seg, hi = lo + seg < n ? lo + seg : n;
  float best = 0.0f;
  uint64_t bi = ~0ull;
  for (uint64_t r = lo + r0; r < hi; r += rps) {
    const float a = fabsf(U[r * k + c]);
    if (a > best) {
      best = a;
      bi = r;
    }
  }
  s_v[tid] = best;
  s_i[tid] = bi;
  __syncthreads();
  if (r0 != 0) return;
  for (uint32_t q = 1; q < rps; ++q) {
    const float v = s_v[q * k + c];
    const uint64_t i = s_i[q * k + c];
    if (v > best || (v == best && v > 0.0f && i < bi)) {
      best = v;
      bi = i;
    }
  }
  pv[(size_t)blockIdx.x * k + c] = best;
  pi[(size_t)blockIdx.x * k + c] = bi;
}
// Pass 2: per column, the partials in block order (earlier rows first).
__global__ void k_pca_sign_pick(const float* __restrict__ U, uint32_t k, uint32_t G,
                                const float* __restrict__ pv, const uint64_t* __restrict__ pi,
                                uint32_t* __restrict__ flip) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  float best = 0.0f;
  uint64_t bi = 0;
  for (uint32_t g = 0; g < G; ++g) {
    const float v = pv[(size_t)g * k + c];
    if (v > best) {
      best = v;
      bi = pi[(size_t)g * k + c];
    }
  }
  flip[c] = U[bi * k + c] < 0.0f;
}
__global__ void k_pca_sign_apply(float* __restrict__ U, uint64_t count, uint32_t k,
                                 const uint32_t* __restrict__ flip) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += step)
    if (flip[t % k]) U[t] = -U[t];
}

// ratio[j] = max(lam_j - sigma, 0) / trace (lam holds eigenvalues of C + sigma I).
__global__ void k_pca_ratio(const double* __restrict__ lam, const double* __restrict__ tv, uint32_t k,
                            float* __restrict__ ratio) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k) ratio[j] = (float)(fmax(lam[j] - tv[1], 0.0) / tv[0]);
}

// ── randomized block Krylov route ──

// Row m of each column c of X (column-major, leading dimension ld) becomes
// -(mu^T X[0, m)[c]): a dense chunk's ones column then subtracts 1 mu^T X from
// its product, the implicit centring (rsvd.rs:252-262). One block per column,
// fixed-order tree sum.
__global__ __launch_bounds__(256) void k_rsvd_mu_row(double* __restrict__ X, uint32_t ld, uint32_t m,
                                                     const float* __restrict__ mu) {
  __shared__ double part[256];
  double* col = X + (size_t)blockIdx.x * ld;
  double a = 0.0;
  for (uint32_t j = threadIdx.x; j < m; j += blockDim.x) a += (double)mu[j] * col[j];
  part[threadIdx.x] = a;
  __syncthreads();
  for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) col[m] = -part[0];
}

// Row m of Z (the chunks' ones column times Y: the column sums 1^T Y) moved to
// zs and cleared, for each of the cols columns.
__global__ void k_rsvd_take_row(double* __restrict__ Z, uint32_t ld, uint32_t m, uint32_t cols,
                                double* __restrict__ zs) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  zs[c] = Z[(size_t)c * ld + m];
  Z[(size_t)c * ld + m] = 0.0;
}

// Z[j][c] -= mu_j zs[c] for j < m: S^T Y - mu (1^T Y) (rsvd.rs:325-335).
__global__ void k_rsvd_center_t(double* __restrict__ Z, uint32_t ld, uint32_t m, uint32_t cols,
                                const float* __restrict__ mu, const double* __restrict__ zs) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x, count = (uint64_t)m * cols;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += step) {
    const uint32_t c = (uint32_t)(t / m), j = (uint32_t)(t - (uint64_t)c * m);
    Z[(size_t)c * ld + j] -= (double)mu[j] * zs[c];
  }
}

// Xr (row-major (m + 1) x cols) = rows 0 .. m of X (column-major, leading
// dimension ld): the layout the sparse product gathers rows from.
__global__ void k_rsvd_rows(const double* __restrict__ X, uint32_t ld, uint32_t m, uint32_t cols,
                            double* __restrict__ Xr) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x, count = (uint64_t)(m + 1) * cols;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += step) {
    const uint64_t j = t / cols;
    Xr[t] = X[(t - j * cols) * ld + j];
  }
}

// Y (column-major n x cols, leading dimension n) = S X + 1 X[m]: the product
// of the CSR matrix with X, whose row m (-mu^T X) centres it (rsvd.rs:241-283).
// One workgroup per 64 rows: wave w takes rows w, w + 4, ... of the tile, lanes
// the columns c0 + lane, each sum over the row's stored entries in order (f64
// fma); the tile goes through LDS and out column by column, lanes over rows.
__global__ __launch_bounds__(256) void k_rsvd_spmm(const uint64_t* __restrict__ indptr,
                                                   const uint32_t* __restrict__ indices,
                                                   const float* __restrict__ val, uint64_t n,
                                                   const double* __restrict__ Xr, uint32_t cols, uint32_t m,
                                                   double* __restrict__ Y) {
  __shared__ double tile[64 * 65];
  const uint64_t row0 = (uint64_t)blockIdx.x * 64;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  for (uint32_t c0 = 0; c0 < cols; c0 += 64) {
    const uint32_t c = c0 + lane;
    const bool cact = c < cols;
    for (uint32_t r = w; r < 64; r += 4) {
      const uint64_t i = row0 + r;
      double acc = 0.0;
      if (i < n && cact) {
        uint64_t e = indptr[i];
        const uint64_t e1 = indptr[i + 1];
        for (; e + 4 <= e1; e += 4) {
          const uint32_t j0 = indices[e], j1 = indices[e + 1], j2 = indices[e + 2], j3 = indices[e + 3];
          const double x0 = Xr[(size_t)j0 * cols + c], x1 = Xr[(size_t)j1 * cols + c];
          const double x2 = Xr[(size_t)j2 * cols + c], x3 = Xr[(size_t)j3 * cols + c];
          acc = fma((double)val[e], x0, acc);
          acc = fma((double)val[e + 1], x1, acc);
          acc = fma((double)val[e + 2], x2, acc);
          acc = fma((double)val[e + 3], x3, acc);
        }
        for (; e < e1; ++e) acc = fma((double)val[e], Xr[(size_t)indices[e] * cols + c], acc);
        acc += Xr[(size_t)m * cols + c];
      }
      tile[r * 65 + lane] = acc;
    }
    __syncthreads();
    const uint64_t i = row0 + lane;
    for (uint32_t cc = w; cc < 64 && c0 + cc < cols; cc += 4)
      if (i < n) Y[(size_t)(c0 + cc) * n + i] = tile[lane * 65 + cc];
    __syncthreads();
  }
}

// G[t] = the np partials' sum, in order.
__global__ void k_pca_fold(const double* __restrict__ Gp, uint32_t np, uint64_t count, double* __restrict__ G) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  double a = 0.0;
  for (uint32_t p = 0; p < np; ++p) a += Gp[(size_t)p * count + t];
  G[t] = a;
}

// Sums of v^2 (f64) over the stored values, per block: block g takes the
// elements g 256 + t + i G 256 in order, then a fixed-order tree sum.
__global__ __launch_bounds__(256) void k_rsvd_sumsq(const float* __restrict__ v, uint64_t nnz,
                                                    double* __restrict__ part) {
  __shared__ double sp[256];
  double a = 0.0;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnz; t += step) {
    const double x = (double)v[t];
    a += x * x;
  }
  sp[threadIdx.x] = a;
  __syncthreads();
  for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sp[threadIdx.x] += sp[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sp[0];
}

// Singular values s_j = sqrt(max(lam_j - sigma, 0)) of the sketch (lam:
// eigenvalues of B B^T + sigma I, descending; sigma = tv[1], 0 without tv),
// their reciprocals (0 for s_j = 0) and the variance ratio
// s_j^2 / total (pca.rs:149-166).
__global__ void k_rsvd_values(const double* __restrict__ lam, const double* __restrict__ tv, uint32_t k,
                              double total, double* __restrict__ sv, double* __restrict__ inv,
                              float* __restrict__ ratio) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const double s = sqrt(fmax(lam[j] - (tv ? tv[1] : 0.0), 0.0));
  sv[j] = s;
  inv[j] = s > 0.0 ? 1.0 / s : 0.0;
  ratio[j] = (float)(s * s / total);
}

// out[i][j] = f32(U[i][j] sc[j]): U column-major (leading dimension ldu), out
// row-major rows x k.
__global__ void k_rsvd_out(const double* __restrict__ U, uint64_t ldu, uint64_t rows, uint32_t k,
                           const double* __restrict__ sc, float* __restrict__ out) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x, count = rows * k;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += step) {
    const uint64_t i = t / k;
    const uint32_t j = (uint32_t)(t - i * k);
    out[t] = (float)(U[(size_t)j * ldu + i] * sc[j]);
  }
}

}  // namespace psq

using namespace psq;

namespace {

unsigned blocks(uint64_t n, unsigned b) { return (unsigned)std::max<uint64_t>(1, (n + b - 1) / b); }
// Grids of the kernels that loop over n- or nnz-sized ranges: a dispatch holds
// fewer than 2^32 work-items, so the grid is capped and the kernel strides.
unsigned stride_blocks(uint64_t n, unsigned b) { return std::min<unsigned>(blocks(n, b), 1u << 20); }

thread_local psq_pca_report g_pca_report;

// Events recorded on the call's stream; destroyed with the call.
struct Events {
  std::vector<hipEvent_t> ev;
  ~Events() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  int record(hipStream_t s, size_t* id) {
    hipEvent_t e;
    PHIPC(hipEventCreate(&e));
    ev.push_back(e);
    *id = ev.size() - 1;
    PHIPC(hipEventRecord(e, s));
    return PSQ_OK;
  }
  double ms(size_t a, size_t b) const {
    float t = 0.0f;
    return hipEventElapsedTime(&t, ev[a], ev[b]) == hipSuccess ? (double)t : 0.0;
  }
};

struct Dev {
  std::vector<void*> ptrs;
  ~Dev() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int alloc(size_t count, T** out) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess)
      return fail(e == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE,
                  "PCA allocation of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    ptrs.push_back(p);
    *out = static_cast<T*>(p);
    return PSQ_OK;
  }
};

// One rocBLAS handle per device for the life of the process (created on first
// use; creation loads the library's kernels).
std::mutex g_blas_mu;
std::map<int, rocblas_handle> g_blas;

int blas_handle(hipStream_t s, rocblas_handle* out) {
  int dev = 0;
  PHIPC(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_blas_mu);
  rocblas_handle& h = g_blas[dev];
  if (!h) {
    PBLAS(rocblas_create_handle(&h));
    PBLAS(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
    PBLAS(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed));  // repeatable sums
  }
  PBLAS(rocblas_set_stream(h, s));
  *out = h;
  return PSQ_OK;
}

size_t small_lds(uint32_t b) { return b <= PCA_LDS_BLOCK ? (size_t)b * b * sizeof(double) : 0; }

// Row segments of the split Gram of a tall m x b block: one (a single DSYRK)
// below 65,536 rows, else one per 4,096 rows, at most 256 and at most 2^25 / b^2
// (the partials' size).
uint32_t gram_parts(uint64_t m, uint32_t b) {
  if (m < 65536) return 1;
  const uint64_t cap = std::max<uint64_t>(2, (1ull << 25) / ((uint64_t)b * b));
  return (uint32_t)std::max<uint64_t>(2, std::min<uint64_t>({m / 4096, 256, cap}));
}

// G (b x b; with parts > 1 full, else its upper triangle) = A^T A for the
// column-major m x b block A. parts > 1: a batched DGEMM per row segment into
// Gp ((parts + 1) b^2 doubles), folded in segment order, so that the product
// of a tall block keeps the GPU busy.
int gram_tall(rocblas_handle h, const double* A, uint32_t m, uint32_t b, double* G, double* Gp, uint32_t parts,
              hipStream_t s) {
  const double one = 1.0, zero = 0.0;
  if (parts <= 1) {
    PBLAS(rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_transpose, b, m, &one, A, m, &zero, G, b));
    return PSQ_OK;
  }
  const uint32_t seg = m / parts, rem = m - seg * parts;
  const rocblas_stride bb = (rocblas_stride)b * b;
  PBLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, b, b, seg, &one, A, m,
                                      seg, A, m, seg, &zero, Gp, b, bb, parts));
  uint32_t np = parts;
  if (rem) {
    const double* Ar = A + (size_t)seg * parts;
    PBLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, b, b, rem, &one, Ar, m, Ar, m, &zero,
                        Gp + (size_t)parts * bb, b));
    ++np;
  }
  hipLaunchKernelGGL(k_pca_fold, dim3(blocks((uint64_t)bb, 256)), dim3(256), 0, s, Gp, np, (uint64_t)bb, G);
  PHIPC(hipGetLastError());
  return PSQ_OK;
}

// Q (m x b, column-major) <- an orthonormal basis of its columns: shifted
// CholeskyQR, then CholeskyQR twice (G = Q^T Q, R = chol(G), Q <- Q R^-1). The
// first j columns keep spanning what the first j input columns span; a column
// dependent on the earlier ones comes out zero (dead: b entries of scratch).
// The Gram is split over `parts` row segments (gram_tall, Gp its partials).
int orthonormalise(rocblas_handle h, double* Q, uint32_t m, uint32_t b, double* G, uint32_t* dead, hipStream_t s,
                   double* Gp = nullptr, uint32_t parts = 1) {
  const double one = 1.0;
  for (int pass = 0; pass < 3; ++pass) {
    PTRY(gram_tall(h, Q, m, b, G, Gp, parts, s));
    hipLaunchKernelGGL(k_pca_chol, dim3(1), dim3(256), small_lds(b), s, G, b, m, pass == 0 ? 1 : 0, dead);
    PHIPC(hipGetLastError());
    PBLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                        rocblas_diagonal_non_unit, m, b, &one, G, b, Q, m));
    if (pass > 0) {
      hipLaunchKernelGGL(k_pca_zero_dead, dim3(std::min(blocks(m, 256), 64u), b), dim3(256), 0, s, Q, m, dead);
      PHIPC(hipGetLastError());
    }
  }
  return PSQ_OK;
}

// ── the Gaussian sketch (rsvd.rs:414-435) ──
// ChaCha with 8 rounds (Bernstein 2008; rand_chacha's ChaCha8Rng: 64-bit block
// counter in words 12-13 from 0, stream 0 in words 14-15, blocks consumed in
// counter order, a u64 as two consecutive words, low first), keyed by
// rand_core's seed_from_u64 (eight PCG32 outputs, little-endian: state <-
// state 6364136223846793005 + 11634580027462260723, then the XSH-RR output);
// a uniform is the u64's top 53 bits times 2^-53, and each Box-Muller pair
// (u1 floored at 1e-300) gives r cos(2 pi u2), r sin(2 pi u2) in f64, rounded
// to f32, filling the rows x cols matrix row by row.
struct ChaCha {
  uint32_t key[8];
  int rounds;
  uint64_t counter = 0;
  uint32_t buf[16];
  int pos = 16;
  static uint32_t rotl(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }
  static void quarter(uint32_t* x, int a, int b, int c, int d) {
    x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
    x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
    x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
    x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
  }
  void block() {
    const uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2],
                             key[3], key[4], key[5], key[6], key[7], (uint32_t)counter,
                             (uint32_t)(counter >> 32), 0u, 0u};
    uint32_t x[16];
    for (int i = 0; i < 16; ++i) x[i] = in[i];
    for (int r = 0; r < rounds; r += 2) {
      quarter(x, 0, 4, 8, 12);
      quarter(x, 1, 5, 9, 13);
      quarter(x, 2, 6, 10, 14);
      quarter(x, 3, 7, 11, 15);
      quarter(x, 0, 5, 10, 15);
      quarter(x, 1, 6, 11, 12);
      quarter(x, 2, 7, 8, 13);
      quarter(x, 3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) buf[i] = x[i] + in[i];
    ++counter;
    pos = 0;
  }
  uint64_t next_u64() {
    if (pos >= 16) block();
    const uint64_t lo = buf[pos], hi = buf[pos + 1];
    pos += 2;
    return lo | (hi << 32);
  }
};

std::vector<float> gaussian_rowmajor(uint64_t rows, uint64_t cols, uint64_t seed) {
  ChaCha g;
  g.rounds = 8;
  uint64_t st = seed;
  for (int i = 0; i < 8; ++i) {
    st = st * 6364136223846793005ull + 11634580027462260723ull;
    const uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27), rot = (uint32_t)(st >> 59);
    g.key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
  }
  const uint64_t count = rows * cols;
  std::vector<float> out(count);
  for (uint64_t i = 0; i < count; i += 2) {
    const double u1 = std::max((double)(g.next_u64() >> 11) * 0x1.0p-53, 1e-300);
    const double u2 = (double)(g.next_u64() >> 11) * 0x1.0p-53;
    const double r = std::sqrt(-2.0 * std::log(u1)), th = 2.0 * M_PI * u2;
    out[i] = (float)(r * std::cos(th));
    if (i + 1 < count) out[i + 1] = (float)(r * std::sin(th));
  }
  return out;
}

// Dense f64 row chunks of the CSR matrix (rows of ld >= m + 1 elements, the
// ones column at m, the rest of the padding zero), one chunk buffer reused.
struct Chunks {
  const uint64_t* ip;
  const uint32_t* ix;
  uint64_t n;
  uint32_t m, ld, R;
  double* D;
  hipStream_t s;
  int densify(uint64_t r0, uint32_t rc, const float* vals) const {
    PHIPC(hipMemsetAsync(D, 0, (size_t)rc * ld * sizeof(double), s));
    hipLaunchKernelGGL(k_pca_densify, dim3(blocks(rc, 4)), dim3(256), 0, s, ip, ix, vals, r0, rc, m, ld, D);
    PHIPC(hipGetLastError());
    return PSQ_OK;
  }
};

// Column sums and sums of squares (f64) of the stored values, over the dense
// chunks, per-segment partials folded in order.
int column_sums(const Chunks& ch, const float* vals, Dev& dv, double* acc_s, double* acc_q) {
  const hipStream_t s = ch.s;
  const uint32_t P = std::min<uint32_t>(PCA_SEGMENTS, ch.R), m = ch.m;
  double *ps_s, *ps_q;
  PTRY(dv.alloc((size_t)P * m, &ps_s));
  PTRY(dv.alloc((size_t)P * m, &ps_q));
  PHIPC(hipMemsetAsync(acc_s, 0, m * sizeof(double), s));
  PHIPC(hipMemsetAsync(acc_q, 0, m * sizeof(double), s));
  for (uint64_t r0 = 0; r0 < ch.n; r0 += ch.R) {
    const uint32_t rc = (uint32_t)std::min<uint64_t>(ch.R, ch.n - r0);
    PTRY(ch.densify(r0, rc, vals));
    const uint32_t Pc = std::min<uint32_t>(P, rc), seg = (rc + Pc - 1) / Pc;
    hipLaunchKernelGGL(k_pca_colsums_part, dim3(blocks(m, 256), Pc), dim3(256), 0, s, ch.D, rc, m, ch.ld, seg,
                       ps_s, ps_q);
    hipLaunchKernelGGL(k_pca_colsums_fold, dim3(blocks(m, 256)), dim3(256), 0, s, ps_s, ps_q, Pc, m, acc_s, acc_q);
    PHIPC(hipGetLastError());
  }
  return PSQ_OK;
}

// Basis width of the randomized route: l = min(k + oversample, min(n, m)) and
// l (n_iter + 1) columns, at most min(n, m) (rsvd.rs:496-499), saturating.
void rsvd_widths(uint64_t n, uint64_t m, uint64_t k, uint64_t oversample, uint64_t n_iter, uint64_t* l_out,
                 uint64_t* width_out) {
  const uint64_t most = std::min(n, m);
  const uint64_t l = std::max<uint64_t>(1, oversample >= most ? most : std::min(k + oversample, most));
  const uint64_t blocks_ = n_iter >= most ? most : n_iter + 1;
  *l_out = l;
  *width_out = blocks_ > most / l ? most : std::min(l * blocks_, most);
}

// Top kk eigenpairs of the symmetric m x m matrix C (column-major, upper
// triangle read; tv = {trace, sigma} on the device, htrace the trace on the
// host): Chebyshev-accelerated subspace iteration on C + sigma I with block
// b = min(m, kk + max(16, kk / 2)). Q (m x b) holds the eigenvectors in its
// first kk columns, descending; lam the eigenvalues of C + sigma I.
int top_eigenpairs(rocblas_handle h, Dev& dv, const double* C, uint32_t m, uint32_t kk, const double* tv,
                   double htrace, psq_pca_report& rep, double** Q_out, double** lam_out, uint32_t* b_out,
                   hipStream_t s) {
  const double one = 1.0, zero = 0.0;
  const uint32_t b = (uint32_t)std::min<uint64_t>(m, kk + std::max<uint32_t>(16, kk / 2));
  if (b > PCA_MAX_BLOCK) return fail(PSQ_E_UNSUPPORTED, "n_components > 680 is not supported");
  double *Q, *Z, *Tmp, *G, *Tb, *Wb, *Ws, *lam, *cs, *res;
  PTRY(dv.alloc((size_t)m * b, &Q));
  PTRY(dv.alloc((size_t)m * b, &Z));
  PTRY(dv.alloc((size_t)m * b, &Tmp));
  PTRY(dv.alloc((size_t)b * b, &G));
  PTRY(dv.alloc((size_t)b * b, &Tb));
  PTRY(dv.alloc((size_t)b * b, &Wb));
  PTRY(dv.alloc((size_t)b * b, &Ws));
  PTRY(dv.alloc(b, &lam));
  PTRY(dv.alloc((size_t)4 * (b / 2 + 1), &cs));
  PTRY(dv.alloc(b, &res));
  uint32_t* dead;
  PTRY(dv.alloc(b, &dead));
  if (htrace <= 0.0) {
    // C = 0 (a PSD matrix with zero trace): no direction is preferred
    hipLaunchKernelGGL(k_pca_unit_block, dim3(blocks((uint64_t)m * kk, 256)), dim3(256), 0, s, Q, lam, m, kk);
    PHIPC(hipGetLastError());
  } else {
    hipLaunchKernelGGL(k_pca_start, dim3(blocks((uint64_t)m * b, 256)), dim3(256), 0, s, Q, m, b);
    PHIPC(hipGetLastError());
    PTRY(orthonormalise(h, Q, m, b, G, dead, s));
    double *Y = nullptr, *H = nullptr;  // the filter's fourth block; lead x b products
    PTRY(dv.alloc((size_t)m * b, &Y));
    PTRY(dv.alloc((size_t)b * b, &H));
    // Zo <- (C + sigma I) X for nb columns, minus V_L diag(lam_L) V_L^T X for the
    // first nl columns of Q (locked Ritz pairs, deflated to eigenvalue 0)
    auto apply = [&](const double* X, double* Zo, uint32_t nb, const double* VL, uint32_t nl) -> int {
      const uint64_t cnt = (uint64_t)m * nb;
      PBLAS(rocblas_dsymm(h, rocblas_side_left, rocblas_fill_upper, m, nb, &one, C, m, X, m, &zero, Zo, m));
      hipLaunchKernelGGL(k_pca_shift, dim3(blocks(cnt, 256)), dim3(256), 0, s, Zo, X, cnt, tv);
      PHIPC(hipGetLastError());
      if (nl) {
        const double minus = -1.0;
        PBLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nl, nb, m, &one, VL, m, X, m,
                            &zero, H, nl));
        hipLaunchKernelGGL(k_pca_scale_rows, dim3(blocks((uint64_t)nl * nb, 256)), dim3(256), 0, s, H, lam, nl,
                           (uint64_t)nl * nb);
        PHIPC(hipGetLastError());
        PBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, nb, nl, &minus, VL, m, H, nl,
                            &one, Zo, m));
      }
      ++rep.iterations;
      return PSQ_OK;
    };
    std::vector<double> hres(kk), hlam(b);
    bool converged = false;
    while (rep.iterations < PCA_MAX_PRODUCTS) {
      PTRY(apply(Q, Z, b, nullptr, 0));
      // Rayleigh-Ritz: T = Q^T Z = W L W^T; Q <- Q W, Z <- Z W
      PBLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, b, b, m, &one, Q, m, Z, m,
                          &zero, Tb, b));
      hipLaunchKernelGGL(k_pca_jacobi, dim3(1), dim3(256), small_lds(b), s, Tb, b, Wb, Ws, lam, cs);
      PHIPC(hipGetLastError());
      PBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, b, b, &one, Q, m, Ws, b, &zero,
                          Tmp, m));
      std::swap(Q, Tmp);
      PBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, b, b, &one, Z, m, Ws, b, &zero,
                          Tmp, m));
      std::swap(Z, Tmp);
      hipLaunchKernelGGL(k_pca_residuals, dim3(kk), dim3(256), 0, s, Z, Q, lam, m, res);
      PHIPC(hipGetLastError());
      PHIPC(hipMemcpyAsync(hres.data(), res, kk * sizeof(double), hipMemcpyDeviceToHost, s));
      PHIPC(hipMemcpyAsync(hlam.data(), lam, b * sizeof(double), hipMemcpyDeviceToHost, s));
      PHIPC(hipStreamSynchronize(s));
      const double top = std::max(std::fabs(hlam[0]), 1e-300);
      converged = true;
      uint32_t lead = kk;  // the converged leading Ritz pairs are locked
      for (uint32_t i = 0; i < kk; ++i)
        if (!(hres[i] <= 1e-11 * top)) {
          converged = false;
          if (lead == kk) lead = i;
        }
      if (m == b) converged = true;  // the whole space: the projection's eigenpairs are C's
      if (converged) break;
      // Chebyshev filter T_d(A / e) on the unlocked columns, A = C + sigma I
      // deflated by the locked pairs, e = the block's smallest Ritz value:
      // components at or below e (the locked ones, now at 0, included) stay
      // bounded by 1, the wanted ones grow by cosh(d acosh(lambda / e)); d keeps
      // the growth of the first unlocked pair under 1e7, so the filtered block
      // stays well conditioned for CholeskyQR.
      const double e = hlam[b - 1];
      const double lu = hlam[lead];
      int d = PCA_MAX_DEGREE;
      if (e > 0.0 && lu > e) d = (int)std::floor(std::acosh(1e7) / std::acosh(lu / e));
      d = std::max(1, std::min(d, PCA_MAX_DEGREE));
      if (!(e > 0.0)) d = 1;
      const uint32_t nb = b - lead;
      const size_t off = (size_t)lead * m;
      // Y0 = Q, Y1 = Z / e, Y_{i+1} = (2 / e) A Y_i - Y_{i-1}, on columns [lead, b)
      // (Z = A Q needs no deflation: the Ritz vectors are orthogonal to the locked ones)
      double *prev = Q, *cur = Z, *nxt = Tmp, *spare = Y;
      hipLaunchKernelGGL(k_pca_axpby, dim3(blocks((uint64_t)m * nb, 256)), dim3(256), 0, s, cur + off, 1.0 / e,
                         cur + off, 0.0, (uint64_t)m * nb);
      PHIPC(hipGetLastError());
      for (int i = 2; i <= d; ++i) {
        PTRY(apply(cur + off, nxt + off, nb, Q, lead));
        hipLaunchKernelGGL(k_pca_axpby, dim3(blocks((uint64_t)m * nb, 256)), dim3(256), 0, s, nxt + off,
                           2.0 / e, prev + off, -1.0, (uint64_t)m * nb);
        PHIPC(hipGetLastError());
        double* t = prev;
        prev = cur;
        cur = nxt;
        if (i == 2) {
          nxt = spare;
          spare = t;  // Q's buffer: the locked columns stay readable there
        } else {
          nxt = t;
        }
      }
      // the new block: the locked columns, then the filtered ones
      if (lead) PHIPC(hipMemcpyAsync(cur, Q, off * sizeof(double), hipMemcpyDeviceToDevice, s));
      double* rest[3];
      int r = 0;
      for (double* p : {Q, Z, Tmp, Y})
        if (p != cur) rest[r++] = p;
      Q = cur;
      Z = rest[0];
      Tmp = rest[1];
      Y = rest[2];
      PTRY(orthonormalise(h, Q, m, b, G, dead, s));
    }
    if (!converged)
      return fail(PSQ_E_DEVICE, "the eigensolver did not converge in %u products", PCA_MAX_PRODUCTS);
  }
  *Q_out = Q;
  *lam_out = lam;
  *b_out = b;
  return PSQ_OK;
}

// Randomized block Krylov SVD of S - 1 mu^T (rsvd.rs:476-566): Q_0 = qr(A' Omega),
// Q_d = qr(A' qr(A'^T Q_{d-1})) for d = 1 .. n_iter while d l < width (a last
// block that does not fit is cut to the width), Q = qr([Q_0 .. Q_q]),
// B^T = A'^T Q; with B B^T = W diag(s^2) W^T, U = Q W[:, :k], embedding = U s,
// columns signed as in the Gram route, ratio = s^2 / ||A'||_F^2, components =
// B^T W[:, :k] / s. A' X is a sparse product over the CSR rows (X staged
// row-major, its row m = -mu^T X the centring); A'^T Y a pass over the dense
// f64 row chunks (DGEMM with the chunk and its ones column, whose row gives
// 1^T Y for the centring). The QR factorisations are shifted CholeskyQR3
// (Grams of tall blocks split over row segments and folded in order) and the
// small eigenproblem the Gram route's solvers, so a run repeats bit for bit. A
// rank-deficient block keeps its dependent columns at zero where Householder QR
// would complete it with unit vectors orthogonal to the matrix's range: either
// way they add nothing to B, and the components of s > 0 are the same.
int rsvd_route(rocblas_handle h, Dev& dv, Events& evs, psq_pca_report& rep, const Chunks& ch, const float* S,
               uint64_t nnz, uint32_t k, int32_t center, uint64_t oversample, uint64_t n_iter, uint64_t seed, float* d_emb,
               float* d_ratio, float* d_comp, size_t ev_start, size_t ev_scaled) {
  const hipStream_t s = ch.s;
  const uint64_t n = ch.n;
  const uint32_t m = ch.m, ld = ch.ld;
  const double one = 1.0, zero = 0.0;
  uint64_t l64, w64;
  rsvd_widths(n, m, k, oversample, n_iter, &l64, &w64);
  const uint32_t l = (uint32_t)l64, width = (uint32_t)w64;

  // 1. column means (f32, rsvd.rs:155-177) and ||A'||_F^2 = sum over stored
  //    values of v^2 - n sum mu^2 (rsvd.rs:186-208), the last sums on the host
  //    in order
  double *acc_s, *acc_q, *part;
  PTRY(dv.alloc(m, &acc_s));
  PTRY(dv.alloc(m, &acc_q));
  PTRY(column_sums(ch, S, dv, acc_s, acc_q));
  ++rep.iterations;
  const uint32_t nnz_blocks = std::min<uint32_t>(1024, blocks(nnz, 256));
  PTRY(dv.alloc(nnz_blocks, &part));
  hipLaunchKernelGGL(k_rsvd_sumsq, dim3(nnz_blocks), dim3(256), 0, s, S, nnz, part);
  PHIPC(hipGetLastError());
  float* mu = nullptr;
  std::vector<float> hmu;
  if (center) {
    PTRY(dv.alloc(m, &mu));
    hipLaunchKernelGGL(k_pca_means, dim3(blocks(m, 256)), dim3(256), 0, s, acc_s, m, n, mu);
    PHIPC(hipGetLastError());
    hmu.resize(m);
    PHIPC(hipMemcpyAsync(hmu.data(), mu, m * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  std::vector<double> hq(nnz_blocks);
  PHIPC(hipMemcpyAsync(hq.data(), part, nnz_blocks * sizeof(double), hipMemcpyDeviceToHost, s));
  PHIPC(hipStreamSynchronize(s));
  double total = 0.0, mu2 = 0.0;
  for (double v : hq) total += v;
  for (float v : hmu) mu2 += (double)v * (double)v;
  total -= (double)n * mu2;
  if (!std::isfinite(total)) return fail(PSQ_E_INVALID, "the matrix holds non-finite values");

  // 2. the Krylov basis: Y = A' X by the sparse product (X_0 = Omega, then
  //    X = qr(A'^T Q_{d-1})), Q_d = qr(Y); then Q = qr(K) and B^T = A'^T Q
  const uint32_t gp = gram_parts(n, width), gz = gram_parts(ld, width), gx = gram_parts(ld, l);
  double *X, *Xr, *Y, *K, *Bt, *G, *Gp, *zs;
  uint32_t* dead;
  PTRY(dv.alloc((size_t)ld * l, &X));
  PTRY(dv.alloc((size_t)(m + 1) * l, &Xr));
  PTRY(dv.alloc((size_t)n * l, &Y));
  PTRY(dv.alloc((size_t)n * width, &K));
  PTRY(dv.alloc((size_t)ld * width, &Bt));
  PTRY(dv.alloc((size_t)width * width, &G));
  PTRY(dv.alloc(std::max<size_t>((size_t)(std::max(gp, gz) + 1) * width * width, (size_t)(gx + 1) * l * l), &Gp));
  PTRY(dv.alloc(width, &zs));
  PTRY(dv.alloc(width, &dead));
  {
    const std::vector<float> om = gaussian_rowmajor(m, l, seed);
    std::vector<double> hx((size_t)ld * l, 0.0);
    for (uint32_t j = 0; j < m; ++j)
      for (uint32_t c = 0; c < l; ++c) hx[(size_t)c * ld + j] = (double)om[(size_t)j * l + c];
    PHIPC(hipMemcpyAsync(X, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice, s));
    PHIPC(hipStreamSynchronize(s));
  }
  std::vector<std::pair<size_t, size_t>> ev_gemm;
  // Y = A' X: X's row m = -mu^T X, X staged row-major, then the sparse product
  auto forward = [&]() -> int {
    if (mu) hipLaunchKernelGGL(k_rsvd_mu_row, dim3(l), dim3(256), 0, s, X, ld, m, mu);
    hipLaunchKernelGGL(k_rsvd_rows, dim3(stride_blocks((uint64_t)(m + 1) * l, 256)), dim3(256), 0, s, X, ld, m, l,
                       Xr);
    hipLaunchKernelGGL(k_rsvd_spmm, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, ch.ip, ch.ix, S, n, Xr, l, m,
                       Y);
    PHIPC(hipGetLastError());
    ++rep.iterations;
    return PSQ_OK;
  };
  // Zo (ld x cols; rows m .. ld - 1 zero) = A'^T Yi (Yi: n x cols, leading dimension n)
  auto backward = [&](const double* Yi, uint32_t cols, double* Zo) -> int {
    for (uint64_t r0 = 0; r0 < n; r0 += ch.R) {
      const uint32_t rc = (uint32_t)std::min<uint64_t>(ch.R, n - r0);
      PTRY(ch.densify(r0, rc, S));
      size_t a, b;
      PTRY(evs.record(s, &a));
      PBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, ld, cols, rc, &one, ch.D, ld, Yi + r0,
                          n, r0 == 0 ? &zero : &one, Zo, ld));
      PTRY(evs.record(s, &b));
      ev_gemm.emplace_back(a, b);
    }
    hipLaunchKernelGGL(k_rsvd_take_row, dim3(blocks(cols, 64)), dim3(64), 0, s, Zo, ld, m, cols, zs);
    if (mu)
      hipLaunchKernelGGL(k_rsvd_center_t, dim3(stride_blocks((uint64_t)m * cols, 256)), dim3(256), 0, s, Zo, ld, m,
                         cols, mu, zs);
    PHIPC(hipGetLastError());
    ++rep.iterations;
    return PSQ_OK;
  };
  PTRY(forward());
  PTRY(orthonormalise(h, Y, (uint32_t)n, l, G, dead, s, Gp, gp));
  PHIPC(hipMemcpyAsync(K, Y, (size_t)n * l * sizeof(double), hipMemcpyDeviceToDevice, s));
  for (uint64_t d = 1; d <= n_iter && d * l < width; ++d) {
    PTRY(backward(Y, l, X));
    PTRY(orthonormalise(h, X, ld, l, G, dead, s, Gp, gx));
    PTRY(forward());
    PTRY(orthonormalise(h, Y, (uint32_t)n, l, G, dead, s, Gp, gp));
    const uint64_t cnt = std::min<uint64_t>(l, width - d * l);
    PHIPC(hipMemcpyAsync(K + d * l * n, Y, (size_t)n * cnt * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  PTRY(orthonormalise(h, K, (uint32_t)n, width, G, dead, s, Gp, gp));
  PTRY(backward(K, width, Bt));
  size_t ev_basis;
  PTRY(evs.record(s, &ev_basis));

  // 3. B B^T = W diag(s^2) W^T (width x width), the top k (descending): the
  //    cyclic Jacobi directly on a small basis, else the Gram route's
  //    eigensolver
  double *Tb, *Wk, *lam, *sv, *inv, *tv = nullptr;
  PTRY(dv.alloc((size_t)width * width, &Tb));
  PTRY(dv.alloc(k, &sv));
  PTRY(dv.alloc(k, &inv));
  PTRY(gram_tall(h, Bt, ld, width, Tb, Gp, gz, s));
  if (width <= PCA_LDS_BLOCK) {
    double *Wb, *cs;
    PTRY(dv.alloc((size_t)width * width, &Wb));
    PTRY(dv.alloc((size_t)width * width, &Wk));
    PTRY(dv.alloc(width, &lam));
    PTRY(dv.alloc((size_t)4 * (width / 2 + 1), &cs));
    if (gz <= 1)  // DSYRK wrote the upper triangle; the Jacobi reads both
      hipLaunchKernelGGL(k_pca_mirror, dim3(blocks((uint64_t)width * width, 256)), dim3(256), 0, s, Tb, width);
    hipLaunchKernelGGL(k_pca_jacobi, dim3(1), dim3(256), small_lds(width), s, Tb, width, Wb, Wk, lam, cs);
    PHIPC(hipGetLastError());
  } else {
    PTRY(dv.alloc(2, &tv));
    hipLaunchKernelGGL(k_pca_trace, dim3(1), dim3(1), 0, s, Tb, width, tv);
    PHIPC(hipGetLastError());
    double htv[2];
    PHIPC(hipMemcpyAsync(htv, tv, sizeof(htv), hipMemcpyDeviceToHost, s));
    PHIPC(hipStreamSynchronize(s));
    psq_pca_report er{};
    uint32_t be;
    PTRY(top_eigenpairs(h, dv, Tb, width, k, tv, htv[0], er, &Wk, &lam, &be, s));
  }
  size_t ev_eigen;
  PTRY(evs.record(s, &ev_eigen));

  // 4. U = Q W[:, :k] (into Y: l >= k columns), the outputs, signs
  PBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, k, width, &one, K,
                      (rocblas_int)n, Wk, width, &zero, Y, (rocblas_int)n));
  hipLaunchKernelGGL(k_rsvd_values, dim3(blocks(k, 64)), dim3(64), 0, s, lam, tv, k, total, sv, inv, d_ratio);
  hipLaunchKernelGGL(k_rsvd_out, dim3(stride_blocks(n * k, 256)), dim3(256), 0, s, Y, n, n, k, sv, d_emb);
  PHIPC(hipGetLastError());
  float* ppv;
  uint64_t* ppi;
  uint32_t* flip;
  const uint32_t rps = std::max<uint32_t>(1, 256 / k);
  const uint32_t Gs = (uint32_t)std::min<uint64_t>(2048, (n + 255) / 256);
  const uint64_t seg = (n + Gs - 1) / Gs;
  PTRY(dv.alloc((size_t)Gs * k, &ppv));
  PTRY(dv.alloc((size_t)Gs * k, &ppi));
  PTRY(dv.alloc(k, &flip));
  hipLaunchKernelGGL(k_pca_sign_part, dim3(Gs), dim3(k * rps), 0, s, d_emb, n, k, rps, seg, ppv, ppi);
  hipLaunchKernelGGL(k_pca_sign_pick, dim3(blocks(k, 64)), dim3(64), 0, s, d_emb, k, Gs, ppv, ppi, flip);
  hipLaunchKernelGGL(k_pca_sign_apply, dim3(stride_blocks(n * k, 256)), dim3(256), 0, s, d_emb, n * k, k, flip);
  PHIPC(hipGetLastError());
  if (d_comp) {
    // V = B^T W[:, :k] / s (into X: ld x l >= ld x k), with the embedding's signs
    PBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, ld, k, width, &one, Bt, ld, Wk, width,
                        &zero, X, ld));
    hipLaunchKernelGGL(k_rsvd_out, dim3(stride_blocks((uint64_t)m * k, 256)), dim3(256), 0, s, X, ld, m, k, inv,
                       d_comp);
    hipLaunchKernelGGL(k_pca_sign_apply, dim3(stride_blocks((uint64_t)m * k, 256)), dim3(256), 0, s, d_comp,
                       (uint64_t)m * k, k, flip);
    PHIPC(hipGetLastError());
  }
  size_t ev_end;
  PTRY(evs.record(s, &ev_end));
  PHIPC(hipStreamSynchronize(s));  // the buffers freed on return are no longer in use
  rep.method = PSQ_PCA_RANDOMIZED;
  rep.block = width;
  rep.rows_per_chunk = ch.R;
  rep.chunks = (n + ch.R - 1) / ch.R;
  rep.ms_scale = evs.ms(ev_start, ev_scaled);
  rep.ms_gram = evs.ms(ev_scaled, ev_basis);
  for (const auto& p : ev_gemm) rep.ms_gemm += evs.ms(p.first, p.second);
  rep.ms_eigen = evs.ms(ev_basis, ev_eigen);
  rep.ms_project = evs.ms(ev_eigen, ev_end);
  rep.ms_total = evs.ms(ev_start, ev_end);
  g_pca_report = rep;
  return PSQ_OK;
}

int pca_device(const uint64_t* ip, const uint32_t* ix, const float* x, uint64_t n, uint64_t m64, uint64_t k64,
               int32_t center, int32_t scale, double max_value, int32_t method, uint64_t oversample,
               uint64_t n_iter, uint64_t seed, float* d_emb, float* d_ratio, float* d_comp, hipStream_t s) {
  const uint32_t m = (uint32_t)m64, kk = (uint32_t)k64;
  const double clip = max_value;  // +inf: no clipping
  if (scale) center = 1;          // pp.scale always centres (pca.rs:416-419)

  uint64_t ends[2] = {0, 0};
  PHIPC(hipMemcpyAsync(&ends[0], ip, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  PHIPC(hipMemcpyAsync(&ends[1], ip + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  PHIPC(hipStreamSynchronize(s));
  if (ends[0] != 0) return fail(PSQ_E_INVALID, "indptr[0] must be 0");
  const uint64_t nnz = ends[1];
  if (nnz > 0 && (!ix || !x)) return fail(PSQ_E_INVALID, "NULL pointer argument");

  rocblas_handle h = nullptr;
  PTRY(blas_handle(s, &h));
  Dev dv;
  Events evs;
  psq_pca_report rep{};
  size_t ev_start, ev_scaled, ev_gram, ev_eigen, ev_end;
  std::vector<std::pair<size_t, size_t>> ev_gemm;
  PTRY(evs.record(s, &ev_start));
  {
    uint32_t* d_bad;
    PTRY(dv.alloc(1, &d_bad));
    PHIPC(hipMemsetAsync(d_bad, 0, 4, s));
    hipLaunchKernelGGL(k_pca_check, dim3(stride_blocks(std::max(nnz, n), 256)), dim3(256), 0, s, ip, n, ix, nnz, m,
                       d_bad);
    PHIPC(hipGetLastError());
    uint32_t bad = 0;
    PHIPC(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, s));
    PHIPC(hipStreamSynchronize(s));
    if (bad & 2) return fail(PSQ_E_INVALID, "indptr must not decrease");
    if (bad & 1) return fail(PSQ_E_INVALID, "a column index is >= n_features");
  }

  // dense f64 row chunks of at least 2^25 elements (256 MB) and 4096 rows, rows
  // of ld = m + 1 (the ones column) rounded up to 8
  // (2^27 elements, 1 GiB, on the randomized route: its passes multiply a chunk
  // by a few columns, so fewer, taller chunks keep the DGEMMs busy)
  const uint32_t ld = (m + 1 + 7) & ~7u;
  const uint64_t chunk_elems = method == PSQ_PCA_RANDOMIZED ? (1ull << 27) : (1ull << 25);
  const uint32_t R = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(n, std::max<uint64_t>(4096, chunk_elems / ld)));
  const uint32_t P = std::min<uint32_t>(PCA_SEGMENTS, R);
  double *D, *acc_s = nullptr, *acc_q = nullptr, *ps_s = nullptr, *ps_q = nullptr, *C, *Gb;
  PTRY(dv.alloc((size_t)R * ld, &D));
  const Chunks ch{ip, ix, n, m, ld, R, D, s};
  const double one = 1.0;
  auto densify = [&](uint64_t r0, uint32_t rc, const float* vals) -> int { return ch.densify(r0, rc, vals); };

  const float* S = x;  // the stored values the Gram and projection read
  if (scale) {
    double *mean, *inv, *zr;
    float* sv;
    PTRY(dv.alloc(m, &acc_s));
    PTRY(dv.alloc(m, &acc_q));
    PTRY(dv.alloc((size_t)P * m, &ps_s));
    PTRY(dv.alloc((size_t)P * m, &ps_q));
    PTRY(dv.alloc(m, &mean));
    PTRY(dv.alloc(m, &inv));
    PTRY(dv.alloc(m, &zr));
    PTRY(dv.alloc(nnz, &sv));
    PHIPC(hipMemsetAsync(acc_s, 0, m * sizeof(double), s));
    PHIPC(hipMemsetAsync(acc_q, 0, m * sizeof(double), s));
    for (uint64_t r0 = 0; r0 < n; r0 += R) {
      const uint32_t rc = (uint32_t)std::min<uint64_t>(R, n - r0);
      PTRY(densify(r0, rc, x));
      const uint32_t Pc = std::min<uint32_t>(P, rc), seg = (rc + Pc - 1) / Pc;
      hipLaunchKernelGGL(k_pca_colsums_part, dim3(blocks(m, 256), Pc), dim3(256), 0, s, D, rc, m, ld, seg, ps_s,
                         ps_q);
      hipLaunchKernelGGL(k_pca_colsums_fold, dim3(blocks(m, 256)), dim3(256), 0, s, ps_s, ps_q, Pc, m, acc_s,
                         acc_q);
      PHIPC(hipGetLastError());
    }
    hipLaunchKernelGGL(k_pca_scale_stats, dim3(blocks(m, 256)), dim3(256), 0, s, acc_s, acc_q, m, n, clip, mean,
                       inv, zr);
    if (nnz)
      hipLaunchKernelGGL(k_pca_scale_values, dim3(stride_blocks(nnz, 256)), dim3(256), 0, s, ix, x, nnz, mean, inv, zr,
                         clip, sv);
    PHIPC(hipGetLastError());
    S = sv;
  }
  PTRY(evs.record(s, &ev_scaled));
  if (method == PSQ_PCA_RANDOMIZED)
    return rsvd_route(h, dv, evs, rep, ch, S, nnz, kk, center, oversample, n_iter, seed, d_emb, d_ratio, d_comp,
                      ev_start, ev_scaled);

  // Gram of the stored values; the ones column's entries hold the column sums
  if (!acc_s) PTRY(dv.alloc(m, &acc_s));
  PTRY(dv.alloc((size_t)m * m, &C));
  PTRY(dv.alloc((size_t)ld * ld, &Gb));
  PHIPC(hipMemsetAsync(Gb, 0, (size_t)ld * ld * sizeof(double), s));
  for (uint64_t r0 = 0; r0 < n; r0 += R) {
    const uint32_t rc = (uint32_t)std::min<uint64_t>(R, n - r0);
    PTRY(densify(r0, rc, S));
    // D is row-major rc x ld, i.e. column-major A = ld x rc: G += A A^T. The full
    // product (DGEMM, N/T) runs near the FP64 matrix peak where DSYRK's half
    // product does not.
    size_t a, b2;
    PTRY(evs.record(s, &a));
    PBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, ld, ld, rc, &one, D, ld, D, ld,
                        &one, Gb, ld));
    PTRY(evs.record(s, &b2));
    ev_gemm.emplace_back(a, b2);
    ++rep.chunks;
  }
  rep.rows_per_chunk = R;
  hipLaunchKernelGGL(k_pca_gram_unpack, dim3(blocks((uint64_t)m * m, 256)), dim3(256), 0, s, Gb, m, ld, C, acc_s);
  PHIPC(hipGetLastError());
  float* mu = nullptr;
  if (center) {
    PTRY(dv.alloc(m, &mu));
    hipLaunchKernelGGL(k_pca_means, dim3(blocks(m, 256)), dim3(256), 0, s, acc_s, m, n, mu);
  }
  hipLaunchKernelGGL(k_pca_center, dim3(blocks((uint64_t)m * m, 256)), dim3(256), 0, s, C, m, n, mu);
  double* tv;
  PTRY(dv.alloc(2, &tv));
  hipLaunchKernelGGL(k_pca_trace, dim3(1), dim3(1), 0, s, C, m, tv);
  PHIPC(hipGetLastError());
  PTRY(evs.record(s, &ev_gram));
  double htv[2];
  PHIPC(hipMemcpyAsync(htv, tv, sizeof(htv), hipMemcpyDeviceToHost, s));
  PHIPC(hipStreamSynchronize(s));
  if (!std::isfinite(htv[0])) return fail(PSQ_E_INVALID, "the matrix holds non-finite values");

  uint32_t b;
  double *Q, *lam;
  PTRY(top_eigenpairs(h, dv, C, m, kk, tv, htv[0], rep, &Q, &lam, &b, s));

  PTRY(evs.record(s, &ev_eigen));
  rep.block = b;

  // projection, signs, ratios
  float *V = d_comp, *mub = nullptr, *ppv;
  uint64_t* ppi;
  uint32_t* flip;
  if (!V) PTRY(dv.alloc((size_t)m * kk, &V));
  hipLaunchKernelGGL(k_pca_vecs, dim3(blocks((uint64_t)m * kk, 256)), dim3(256), 0, s, Q, m, kk, V);
  if (center) {
    PTRY(dv.alloc(kk, &mub));
    hipLaunchKernelGGL(k_pca_mub, dim3(blocks(kk, 64)), dim3(64), 0, s, mu, V, m, kk, mub);
  }
  hipLaunchKernelGGL(k_pca_project, dim3(stride_blocks(n * kk, 256)), dim3(256), 0, s, ip, ix, S, n, V, kk, mub, d_emb);
  const uint32_t rps = std::max<uint32_t>(1, 256 / kk);
  const uint32_t Gs = (uint32_t)std::min<uint64_t>(2048, (n + 255) / 256);
  const uint64_t seg = (n + Gs - 1) / Gs;
  PTRY(dv.alloc((size_t)Gs * kk, &ppv));
  PTRY(dv.alloc((size_t)Gs * kk, &ppi));
  PTRY(dv.alloc(kk, &flip));
  hipLaunchKernelGGL(k_pca_sign_part, dim3(Gs), dim3(kk * rps), 0, s, d_emb, n, kk, rps, seg, ppv, ppi);
  hipLaunchKernelGGL(k_pca_sign_pick, dim3(blocks(kk, 64)), dim3(64), 0, s, d_emb, kk, Gs, ppv, ppi, flip);
  hipLaunchKernelGGL(k_pca_sign_apply, dim3(stride_blocks(n * kk, 256)), dim3(256), 0, s, d_emb, n * kk, kk, flip);
  // the axes take the embedding's column signs: (X - mu) V reproduces the embedding
  if (d_comp)
    hipLaunchKernelGGL(k_pca_sign_apply, dim3(stride_blocks((uint64_t)m * kk, 256)), dim3(256), 0, s, d_comp,
                       (uint64_t)m * kk, kk, flip);
  hipLaunchKernelGGL(k_pca_ratio, dim3(blocks(kk, 64)), dim3(64), 0, s, lam, tv, kk, d_ratio);
  PHIPC(hipGetLastError());
  PTRY(evs.record(s, &ev_end));
  PHIPC(hipStreamSynchronize(s));  // the buffers freed on return are no longer in use
  rep.ms_scale = evs.ms(ev_start, ev_scaled);
  rep.ms_gram = evs.ms(ev_scaled, ev_gram);
  for (const auto& p : ev_gemm) rep.ms_gemm += evs.ms(p.first, p.second);
  rep.ms_eigen = evs.ms(ev_gram, ev_eigen);
  rep.ms_project = evs.ms(ev_eigen, ev_end);
  rep.ms_total = evs.ms(ev_start, ev_end);
  rep.method = PSQ_PCA_GRAM;
  g_pca_report = rep;
  return PSQ_OK;
}

// Resolves PSQ_PCA_AUTO in *method: the Gram route up to PCA_MAX_FEATURES genes,
// the randomized route above.
int pca_validate(uint64_t n, uint64_t m, uint64_t k, int32_t scale, double max_value, int32_t* method,
                 uint64_t oversample, uint64_t n_iter) {
  if (n == 0 || m == 0)
    return fail(PSQ_E_INVALID, "input matrix has zero rows (%llu) or columns (%llu)", (unsigned long long)n,
                (unsigned long long)m);
  const uint64_t most = std::min(n, m);
  if (k > most)
    return fail(PSQ_E_INVALID, "n_components (%llu) exceeds min(n_obs, n_features) (%llu)",
                (unsigned long long)k, (unsigned long long)most);
  if (k == 0) return fail(PSQ_E_INVALID, "n_components must be >= 1");
  if (scale && !(max_value >= 0.0)) return fail(PSQ_E_INVALID, "max_value must be >= 0 (or +inf for none)");
  if (n >= (1ull << 40) / k) return fail(PSQ_E_UNSUPPORTED, "n_obs * n_components must be < 2^40");
  if (*method != PSQ_PCA_AUTO && *method != PSQ_PCA_GRAM && *method != PSQ_PCA_RANDOMIZED)
    return fail(PSQ_E_INVALID, "method must be 0 (auto), 1 (Gram) or 2 (randomized), got %d", *method);
  if (*method == PSQ_PCA_AUTO) *method = m > PCA_MAX_FEATURES ? PSQ_PCA_RANDOMIZED : PSQ_PCA_GRAM;
  if (*method == PSQ_PCA_GRAM) {
    if (m > PCA_MAX_FEATURES)
      return fail(PSQ_E_UNSUPPORTED, "n_features (%llu) > %u is not supported by the Gram route",
                  (unsigned long long)m, PCA_MAX_FEATURES);
    return PSQ_OK;
  }
  // rocBLAS dimensions are 32-bit: n_obs rows and ld = n_features + 1 rounded to 8
  if (n > 0x7FFFFFFFull || m > 0x7FFFFFF0ull)
    return fail(PSQ_E_UNSUPPORTED, "the randomized route takes at most 2^31 - 1 cells and 2^31 - 16 genes");
  uint64_t l, width;
  rsvd_widths(n, m, k, oversample, n_iter, &l, &width);
  if (width > PCA_MAX_BLOCK)
    return fail(PSQ_E_UNSUPPORTED,
                "the Krylov basis of %llu columns ((n_components + oversample) (n_iter + 1)) exceeds %u",
                (unsigned long long)width, PCA_MAX_BLOCK);
  return PSQ_OK;
}

}  // namespace

extern "C" int psq_pca_csr_device_ex(const uint64_t* d_indptr, const uint32_t* d_indices, const float* d_data,
                                     uint64_t n_obs, uint64_t n_features, uint64_t n_components, int32_t center,
                                     int32_t scale, double max_value, int32_t method, uint64_t oversample,
                                     uint64_t n_iter, uint64_t seed, float* d_embedding, float* d_variance_ratio,
                                     float* d_components, void* stream) {
  PTRY(pca_validate(n_obs, n_features, n_components, scale, max_value, &method, oversample, n_iter));
  if (!d_indptr || !d_embedding || !d_variance_ratio) return fail(PSQ_E_INVALID, "NULL pointer argument");
  return pca_device(d_indptr, d_indices, d_data, n_obs, n_features, n_components, center, scale, max_value,
                    method, oversample, n_iter, seed, d_embedding, d_variance_ratio, d_components,
                    (hipStream_t)stream);
}

extern "C" int psq_pca_csr_ex(const uint64_t* indptr, const uint32_t* indices, const float* data, uint64_t n_obs,
                              uint64_t n_features, uint64_t n_components, int32_t center, int32_t scale,
                              double max_value, int32_t method, uint64_t oversample, uint64_t n_iter,
                              uint64_t seed, float* embedding, float* variance_ratio, float* components,
                              int32_t device) {
  PTRY(pca_validate(n_obs, n_features, n_components, scale, max_value, &method, oversample, n_iter));
  if (!indptr || !embedding || !variance_ratio) return fail(PSQ_E_INVALID, "NULL pointer argument");
  const uint64_t n = n_obs, m = n_features, k = n_components, nnz = indptr[n];
  if (nnz > 0 && (!indices || !data)) return fail(PSQ_E_INVALID, "NULL pointer argument");
  int count = 0;
  PHIPC(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(PSQ_E_INVALID, "device %d out of range", device);
  PHIPC(hipSetDevice(device));
  hipStream_t s = nullptr;
  Dev dv;
  uint64_t* d_ip;
  uint32_t* d_ix;
  float *d_x, *d_emb, *d_ratio, *d_comp = nullptr;
  PTRY(dv.alloc(n + 1, &d_ip));
  PTRY(dv.alloc(nnz, &d_ix));
  PTRY(dv.alloc(nnz, &d_x));
  PTRY(dv.alloc(n * k, &d_emb));
  PTRY(dv.alloc(k, &d_ratio));
  if (components) PTRY(dv.alloc(m * k, &d_comp));
  PHIPC(hipMemcpyAsync(d_ip, indptr, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  if (nnz) {
    PHIPC(hipMemcpyAsync(d_ix, indices, nnz * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    PHIPC(hipMemcpyAsync(d_x, data, nnz * sizeof(float), hipMemcpyHostToDevice, s));
  }
  PTRY(pca_device(d_ip, d_ix, d_x, n, m, k, center, scale, max_value, method, oversample, n_iter, seed, d_emb,
                  d_ratio, d_comp, s));
  PHIPC(hipMemcpyAsync(embedding, d_emb, n * k * sizeof(float), hipMemcpyDeviceToHost, s));
  PHIPC(hipMemcpyAsync(variance_ratio, d_ratio, k * sizeof(float), hipMemcpyDeviceToHost, s));
  if (components) PHIPC(hipMemcpyAsync(components, d_comp, m * k * sizeof(float), hipMemcpyDeviceToHost, s));
  PHIPC(hipStreamSynchronize(s));
  return PSQ_OK;
}

// The reference's default PcaOptions (pca.rs:73-83): oversample 10, n_iter 1, seed 0.
extern "C" int psq_pca_csr_device(const uint64_t* d_indptr, const uint32_t* d_indices, const float* d_data,
                                  uint64_t n_obs, uint64_t n_features, uint64_t n_components, int32_t center,
                                  int32_t scale, double max_value, float* d_embedding, float* d_variance_ratio,
                                  float* d_components, void* stream) {
  return psq_pca_csr_device_ex(d_indptr, d_indices, d_data, n_obs, n_features, n_components, center, scale,
                               max_value, PSQ_PCA_AUTO, 10, 1, 0, d_embedding, d_variance_ratio, d_components,
                               stream);
}

extern "C" int psq_pca_csr(const uint64_t* indptr, const uint32_t* indices, const float* data, uint64_t n_obs,
                           uint64_t n_features, uint64_t n_components, int32_t center, int32_t scale,
                           double max_value, float* embedding, float* variance_ratio, float* components,
                           int32_t device) {
  return psq_pca_csr_ex(indptr, indices, data, n_obs, n_features, n_components, center, scale, max_value,
                        PSQ_PCA_AUTO, 10, 1, 0, embedding, variance_ratio, components, device);
}

extern "C" int psq_pca_last_report(psq_pca_report* out) {
  if (!out) return fail(PSQ_E_INVALID, "NULL pointer argument");
  *out = g_pca_report;
  return PSQ_OK;
}
