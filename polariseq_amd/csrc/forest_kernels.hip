// Random-projection forest start for NN-Descent (gfx950).
//
// Reference: crates/polariseq-core/src/nndescent.rs:356-641 (trees) and 767-817
// (all pairs inside each leaf, applied tree after tree). The GPU builds every
// tree at once, level by level. Per tree, `ids` is a permutation of the points
// and every node is a contiguous segment of it. At level L, a node (start, c)
// with c > leaf_size:
//   * takes the hyperplane through members a = seg[h(L, start, 0) % c] and
//     b = seg[h(L, start, r) % c] (r = 1.., up to 8 redraws while b == a):
//     normal = x_a - x_b, offset = -Σ normal_t * 0.5 * (x_a,t + x_b,t), a
//     sequential f32 fold (nndescent.rs:367-385);
//   * sends a point left when dot(normal, x) + offset > 0 (the 8-lane dot of
//     simd.rs:22-37); |margin| < 1e-8 draws a hash bit (nndescent.rs:439-444);
//   * if every point lands on one side, hash bits split it instead, and if
//     that fails too the node becomes a leaf;
//   * is partitioned stably: left members first, in their order.
// Children with at most leaf_size points (or past depth 100) become leaves.
// Then, tree by tree, each point merges the distances to its leaf-mates into
// its row (top-k, deduplicated). oracle/gpu_model.cpp states the same steps.
#include "psq_kernels.hpp"

namespace psq {

constexpr uint32_t NONE = 0xFFFFFFFFu;

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

// FP16 copy of the split rows (dh halves per row, zero padded), read by the
// margin screen below: one 64- or 128-byte line per row instead of two.
__global__ void k_rows_to_half(const float* __restrict__ X, uint32_t d, uint32_t dp, uint32_t n,
                               _Float16* __restrict__ Xh, uint32_t dh) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * dh) return;
  const size_t r = e / dh;
  const uint32_t j = (uint32_t)(e - r * dh);
  Xh[e] = (_Float16)(j < d ? X[r * dp + j] : 0.0f);
}

// Which side of the hyperplane (normal nv, offset off, nabs = sum |nv_j|) a
// point lies on. The reference decides on m = dot_ref(nv, x) + off (f32,
// simd.rs:22-37 order; |m| < 1e-8 takes hash bit 2, nndescent.rs:439-444).
// Screen: m~ = sum nv_j * half(x_j) + off from the FP16 copy, A = sum |nv_j *
// half(x_j)|. FP16 rounding moves each term by at most 2^-11 of itself (2^-25
// absolutely for subnormals), and both f32 sums are within d * 2^-24 * A of
// their real values, so |m~ - m| <= 1e-3 A + 1e-7 nabs + 1e-6 |m~| =: E (a
// factor 2 to spare at d <= 256). When |m~| > E + 1e-8, m has the sign of m~
// and is not a tie: returns 1 (left) or 0. Otherwise returns 2: the exact
// margin (forest_side_exact) decides, in a separate pass so that the few
// points near a plane do not hold up whole waves.
__device__ inline uint32_t forest_side_screen(const float* nv, float off, float nabs,
                                              const _Float16* __restrict__ xh, uint32_t d) {
  // nv and xh are 16-byte aligned; chunks of 8 with 8 independent sums
  float m[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const uint32_t full = d >> 3;
  for (uint32_t c = 0; c < full; ++c) {
    const half8 h = *reinterpret_cast<const half8*>(xh + 8 * c);
    const float4 n0 = *reinterpret_cast<const float4*>(nv + 8 * c);
    const float4 n1 = *reinterpret_cast<const float4*>(nv + 8 * c + 4);
    const float p[8] = {n0.x * (float)h[0], n0.y * (float)h[1], n0.z * (float)h[2],
                        n0.w * (float)h[3], n1.x * (float)h[4], n1.y * (float)h[5],
                        n1.z * (float)h[6], n1.w * (float)h[7]};
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      m[u] += p[u];
      a[u] += fabsf(p[u]);
    }
  }
  float mt = ((m[0] + m[4]) + (m[1] + m[5])) + ((m[2] + m[6]) + (m[3] + m[7]));
  float A = ((a[0] + a[4]) + (a[1] + a[5])) + ((a[2] + a[6]) + (a[3] + a[7]));
  for (uint32_t j = full * 8; j < d; ++j) {
    const float p = nv[j] * (float)xh[j];
    mt += p;
    A += fabsf(p);
  }
  mt += off;
  if (fabsf(mt) > 1e-3f * A + 1e-7f * nabs + 1e-6f * fabsf(mt) + 1e-8f) return mt > 0.0f ? 1u : 0u;
  return 2u;
}
__device__ inline uint32_t forest_side_exact(const float* nv, float off, const float* __restrict__ x,
                                             uint32_t d, uint64_t ts, uint32_t level, uint32_t g) {
  const float margin = dot_ref(nv, x, (int)d) + off;
  if (fabsf(margin) < 1e-8f) return (tree_hash(ts, level, g, 2) & 1) == 0 ? 1u : 0u;
  return margin > 0.0f ? 1u : 0u;
}

__global__ void k_forest_init(uint32_t* __restrict__ ids, uint32_t n, uint32_t T) {
  size_t pos = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= (size_t)n * T) return;
  uint32_t t = (uint32_t)(pos / n);
  ids[pos] = (uint32_t)(pos - (size_t)t * n);
}

// One wave per active node: pivots, normal and offset.
__global__ __launch_bounds__(256) void k_forest_planes(
    const float* __restrict__ Xp, uint32_t d, uint32_t dp, uint32_t n,
    const uint32_t* __restrict__ ids, const uint32_t* __restrict__ nstart,
    const uint32_t* __restrict__ ncount, const uint32_t* __restrict__ ntree, uint32_t A,
    const uint64_t* __restrict__ tseeds, uint32_t level, float* __restrict__ normals,
    float* __restrict__ offsets, float* __restrict__ nabs) {
  const int w = threadIdx.x / WAVE, l = lane_id();
  const uint32_t v = blockIdx.x * 4 + w;
  if (v >= A) return;
  const uint32_t start = nstart[v], c = ncount[v], t = ntree[v];
  const uint64_t ts = tseeds[t];
  const uint32_t* seg = ids + (size_t)t * n + start;
  const uint32_t a = seg[tree_hash(ts, level, start, 0) % c];
  uint32_t b = seg[tree_hash(ts, level, start, 1) % c];
  for (uint64_t r = 2; b == a && r < 10; ++r) b = seg[tree_hash(ts, level, start, r) % c];
  const float* xa = Xp + (size_t)a * dp;
  const float* xb = Xp + (size_t)b * dp;
  float* nv = normals + (size_t)v * dp;
  for (uint32_t j = l; j < dp; j += WAVE) nv[j] = j < d ? xa[j] - xb[j] : 0.0f;
  if (l == 0) {
    float s = -0.0f, na = 0.0f;
    for (uint32_t j = 0; j < d; ++j) {
      float nj = xa[j] - xb[j];
      s = s + nj * 0.5f * (xa[j] + xb[j]);
      na += fabsf(nj);
    }
    offsets[v] = -s;
    nabs[v] = na;
  }
}

// ── level passes over the positions of active nodes only ──
// An active node (start, count, tree) is cut into tiles of FT_TILE positions;
// toff (exclusive scan of the tile counts) numbers the tiles node after node.
// Per tile: the side bits of its positions (hyperplane side in fl, hash
// fallback in fr) as bit masks in position order, and their counts, whose
// scans over the tiles give every tile's rank base inside its node.

// Tiles per active node.
__global__ void k_ftile_count(const uint32_t* __restrict__ ncount, uint32_t A,
                              uint32_t* __restrict__ ntile) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < A) ntile[v] = (ncount[v] + FT_TILE - 1) / FT_TILE;
}
// Tile -> node (binary search in toff[0..A]).
__global__ void k_ftile_map(const uint32_t* __restrict__ toff, uint32_t A, uint32_t NT,
                            uint32_t* __restrict__ tnode) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NT) return;
  uint32_t lo = 0, hi = A;  // toff[lo] <= t < toff[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (toff[mid] <= t) lo = mid; else hi = mid;
  }
  tnode[t] = lo;
}

// One block of 256 per tile, position e = tid + 256 u of the tile (u < 4).
__global__ __launch_bounds__(256) void k_forest_sides_tile(
    const float* __restrict__ Xp, uint32_t d, uint32_t dp, uint32_t n,
    const _Float16* __restrict__ Xh, uint32_t dh, const uint32_t* __restrict__ ids,
    const uint32_t* __restrict__ nstart, const uint32_t* __restrict__ ncount,
    const uint32_t* __restrict__ ntree, const uint32_t* __restrict__ toff,
    const uint32_t* __restrict__ tnode, const uint64_t* __restrict__ tseeds, uint32_t level,
    const float* __restrict__ normals, const float* __restrict__ offsets,
    const float* __restrict__ nabs, uint64_t* __restrict__ flmask, uint64_t* __restrict__ frmask,
    uint32_t* __restrict__ tcl, uint32_t* __restrict__ tcr) {
  __shared__ uint32_t s_cl[4], s_cr[4];
  const uint32_t t = blockIdx.x, tid = threadIdx.x, w = tid / WAVE;
  const uint32_t v = tnode[t];
  const uint32_t s0 = nstart[v], c = ncount[v], tr = ntree[v];
  const uint32_t e0 = (t - toff[v]) * FT_TILE;
  const uint64_t ts = tseeds[tr];
  const float* nv = normals + (size_t)v * dp;
  const float off = offsets[v], na = nabs[v];
  const uint32_t* tids = ids + (size_t)tr * n + s0;
  uint32_t cl = 0, cr = 0;
#pragma unroll
  for (int u = 0; u < FT_TILE / 256; ++u) {
    const uint32_t e = tid + 256 * u, pin = e0 + e;
    uint32_t fl = 0, fr = 0;
    if (pin < c) {
      const uint32_t i = tids[pin];
      fl = forest_side_screen(nv, off, na, Xh + (size_t)i * dh, d);
      if (fl == 2u) fl = forest_side_exact(nv, off, Xp + (size_t)i * dp, d, ts, level, i);
      fr = (tree_hash(ts, level, i, 3) & 1) == 0 ? 1u : 0u;
    }
    const uint64_t ml = __ballot(fl != 0), mr = __ballot(fr != 0);
    if (lane_id() == 0) {
      flmask[(size_t)t * (FT_TILE / 64) + e / 64] = ml;
      frmask[(size_t)t * (FT_TILE / 64) + e / 64] = mr;
    }
    cl += (uint32_t)__popcll(ml);
    cr += (uint32_t)__popcll(mr);
  }
  if (lane_id() == 0) {
    s_cl[w] = cl;
    s_cr[w] = cr;
  }
  __syncthreads();
  if (tid == 0) {
    tcl[t] = (s_cl[0] + s_cl[1]) + (s_cl[2] + s_cl[3]);
    tcr[t] = (s_cr[0] + s_cr[1]) + (s_cr[2] + s_cr[3]);
  }
}

// Split mode per active node from the tile scans: mode 0 = hyperplane, 1 =
// hash bits, 2 = no split (a leaf); and how many active / small children.
__global__ void k_forest_nodes_tile(const uint32_t* __restrict__ ncount, uint32_t A,
                                    const uint32_t* __restrict__ toff,
                                    const uint32_t* __restrict__ scl, const uint32_t* __restrict__ scr,
                                    uint32_t leaf_size, uint32_t cap, uint32_t level,
                                    uint32_t* __restrict__ nl_out, uint32_t* __restrict__ mode_out,
                                    uint32_t* __restrict__ nact, uint32_t* __restrict__ nsmall) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= A) return;
  const uint32_t c = ncount[v];
  const uint32_t cl = scl[toff[v + 1]] - scl[toff[v]];
  const uint32_t rl = scr[toff[v + 1]] - scr[toff[v]];
  uint32_t mode, nl;
  if (cl != 0 && cl != c) {
    mode = 0;
    nl = cl;
  } else if (rl != 0 && rl != c) {
    mode = 1;
    nl = rl;
  } else {
    mode = 2;
    nl = c;
  }
  mode_out[v] = mode;
  nl_out[v] = nl;
  if (mode == 2) {
    nact[v] = 0;
    nsmall[v] = 0;
    return;
  }
  const bool child_deep = level + 1 > 100;
  uint32_t a = 0, sm = 0;
  const uint32_t cc[2] = {nl, c - nl};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (child_deep || cc[q] <= leaf_size) continue;
    if (cc[q] <= cap) ++sm;
    else ++a;
  }
  nact[v] = a;
  nsmall[v] = sm;
}

// Stable partition of each tile's positions into tmp (same tree offsets as
// ids): left members first, in position order. One block of 256 per tile.
__global__ __launch_bounds__(256) void k_forest_scatter_tile(
    uint32_t n, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ nstart,
    const uint32_t* __restrict__ ncount, const uint32_t* __restrict__ ntree,
    const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tnode,
    const uint32_t* __restrict__ nl_in, const uint32_t* __restrict__ mode_in,
    const uint64_t* __restrict__ flmask, const uint64_t* __restrict__ frmask,
    const uint32_t* __restrict__ scl, const uint32_t* __restrict__ scr, uint32_t* __restrict__ tmp) {
  constexpr uint32_t W = FT_TILE / 64;
  __shared__ uint64_t s_m[W];
  __shared__ uint32_t s_pre[W];
  const uint32_t t = blockIdx.x, tid = threadIdx.x;
  const uint32_t v = tnode[t];
  const uint32_t s0 = nstart[v], c = ncount[v], tr = ntree[v];
  const uint32_t e0 = (t - toff[v]) * FT_TILE;
  const uint32_t mode = mode_in[v];
  const uint32_t* X = mode == 1 ? scr : scl;
  const uint64_t* M = mode == 1 ? frmask : flmask;
  if (tid < W) s_m[tid] = M[(size_t)t * W + tid];
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (uint32_t q = 0; q < W; ++q) {
      s_pre[q] = acc;
      acc += (uint32_t)__popcll(s_m[q]);
    }
  }
  __syncthreads();
  const uint32_t base = X[t] - X[toff[v]];  // left members of the node before this tile
  const uint32_t nl = nl_in[v];
  const uint32_t* tids = ids + (size_t)tr * n + s0;
  uint32_t* out = tmp + (size_t)tr * n + s0;
#pragma unroll
  for (int u = 0; u < FT_TILE / 256; ++u) {
    const uint32_t e = tid + 256 * u, pin = e0 + e;
    if (pin >= c) continue;
    const uint32_t id = tids[pin];
    if (mode == 2) {
      out[pin] = id;
      continue;
    }
    const uint64_t word = s_m[e / 64];
    const uint32_t b = e % 64;
    const uint32_t rank = base + s_pre[e / 64] + (uint32_t)__popcll(word & ((1ull << b) - 1ull));
    const bool left = (word >> b) & 1ull;
    out[left ? rank : nl + (pin - rank)] = id;
  }
}

// tmp -> ids for the positions of every tile.
__global__ __launch_bounds__(256) void k_forest_copyback_tile(
    uint32_t n, const uint32_t* __restrict__ nstart, const uint32_t* __restrict__ ncount,
    const uint32_t* __restrict__ ntree, const uint32_t* __restrict__ toff,
    const uint32_t* __restrict__ tnode, const uint32_t* __restrict__ tmp, uint32_t* __restrict__ ids) {
  const uint32_t t = blockIdx.x, tid = threadIdx.x;
  const uint32_t v = tnode[t];
  const uint32_t s0 = nstart[v], c = ncount[v], tr = ntree[v];
  const uint32_t e0 = (t - toff[v]) * FT_TILE;
  const size_t g0 = (size_t)tr * n + s0;
#pragma unroll
  for (int u = 0; u < FT_TILE / 256; ++u) {
    const uint32_t pin = e0 + tid + 256 * u;
    if (pin < c) ids[g0 + pin] = tmp[g0 + pin];
  }
}

// One thread per active node: next-level nodes, small subtrees, leaves (the
// count at the leaf's start position) and the child map.
__global__ void k_forest_emit(const uint32_t* __restrict__ nstart, const uint32_t* __restrict__ ncount,
                              const uint32_t* __restrict__ ntree, uint32_t A, uint32_t n,
                              const uint32_t* __restrict__ nl_in, const uint32_t* __restrict__ mode_in,
                              const uint32_t* __restrict__ act_off, const uint32_t* __restrict__ small_off,
                              uint32_t leaf_size, uint32_t cap, uint32_t level,
                              uint32_t* __restrict__ x_start, uint32_t* __restrict__ x_count,
                              uint32_t* __restrict__ x_tree, uint32_t* __restrict__ child,
                              uint32_t* __restrict__ lcnt_pos, uint32_t small_base,
                              uint32_t* __restrict__ sm_start, uint32_t* __restrict__ sm_count,
                              uint32_t* __restrict__ sm_tree, uint32_t* __restrict__ sm_level) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= A) return;
  const uint32_t start = nstart[v], c = ncount[v], t = ntree[v];
  const uint32_t mode = mode_in[v], nl = nl_in[v];
  uint32_t ao = act_off[v], so = small_base + small_off[v];
  if (mode == 2) {
    lcnt_pos[(size_t)t * n + start] = c;
    child[2 * v] = NONE;
    child[2 * v + 1] = NONE;
    return;
  }
  const bool child_deep = level + 1 > 100;
  const uint32_t cs[2] = {start, start + nl};
  const uint32_t cc[2] = {nl, c - nl};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    if (child_deep || cc[s] <= leaf_size) {
      lcnt_pos[(size_t)t * n + cs[s]] = cc[s];
      child[2 * v + s] = NONE;
    } else if (cc[s] <= cap) {
      sm_start[so] = cs[s];
      sm_count[so] = cc[s];
      sm_tree[so] = t;
      sm_level[so] = level + 1;
      ++so;
      child[2 * v + s] = NONE;
    } else {
      x_start[ao] = cs[s];
      x_count[ao] = cc[s];
      x_tree[ao] = t;
      child[2 * v + s] = ao;
      ++ao;
    }
  }
}

// dot_ref (simd.rs:22-37 order) of an LDS normal with a global row whose
// chunks of 8 are read as two 16-byte loads (rows are 16-byte aligned).
__device__ inline float dot_ref_row4(const float* nv, const float* __restrict__ row, uint32_t d) {
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const uint32_t c = d >> 3;
  for (uint32_t ch = 0; ch < c; ++ch) {
    const float4 a = *reinterpret_cast<const float4*>(row + ch * 8);
    const float4 b = *reinterpret_cast<const float4*>(row + ch * 8 + 4);
    acc[0] = acc[0] + nv[ch * 8 + 0] * a.x;
    acc[1] = acc[1] + nv[ch * 8 + 1] * a.y;
    acc[2] = acc[2] + nv[ch * 8 + 2] * a.z;
    acc[3] = acc[3] + nv[ch * 8 + 3] * a.w;
    acc[4] = acc[4] + nv[ch * 8 + 4] * b.x;
    acc[5] = acc[5] + nv[ch * 8 + 5] * b.y;
    acc[6] = acc[6] + nv[ch * 8 + 6] * b.z;
    acc[7] = acc[7] + nv[ch * 8 + 7] * b.w;
  }
  float tail = 0.f;
  for (uint32_t i = c * 8; i < d; ++i) tail = tail + nv[i] * row[i];
  return (((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]))) + tail;
}

// Subtrees of at most C (<= 256) points are finished by one wave each (a block
// of 64 threads, so __syncthreads() is a wave barrier): every remaining level
// of the subtree runs with the steps of the level kernels above (same hashes
// on the global segment start, hyperplanes, offset fold, margins, tie and
// fallback rules, stable partition), so the tree equals the one the level
// kernels would build. The subtree's rows, read once per level, stay in the
// cache; all bookkeeping is in LDS. Leaves are written as counts at their
// start positions; the final order of the subtree's positions goes back to ids.
// LDS (u32 unless noted): normals[NC][RSs] f32 | terms[NC][RSs] f32 |
// offsets[NC] f32 | gid[C] | lid[2][C] | pn[2][C] | scan_l[C+1] | scan_r[C+1] |
// nst[2][NC] | ncn[2][NC] | mode[NC] | nl[NC] | child[2 NC] | piv_a[NC] |
// piv_b[NC]. NC bounds the active nodes of one level (C / (leaf_size + 1) + 1).
__global__ __launch_bounds__(64) void k_forest_finish(
    const float* __restrict__ Xs, uint32_t ds, uint32_t dps, uint32_t RSs, uint32_t n,
    uint32_t* __restrict__ ids, const uint32_t* __restrict__ sm_start,
    const uint32_t* __restrict__ sm_count, const uint32_t* __restrict__ sm_tree,
    const uint32_t* __restrict__ sm_level, const uint64_t* __restrict__ tseeds, uint32_t leaf_size,
    uint32_t C, uint32_t NC, uint32_t* __restrict__ lcnt_pos, uint32_t nsmall) {
  extern __shared__ __attribute__((aligned(16))) float ldsf[];
  __shared__ uint32_t s_nA;
  float* normals = ldsf;                            // NC * RSs, 16-byte aligned rows
  float* terms = normals + (size_t)NC * RSs;        // NC * RSs
  float* offs = terms + (size_t)NC * RSs;           // NC
  uint32_t* gid = reinterpret_cast<uint32_t*>(offs + NC);
  uint32_t* lid[2] = {gid + C, gid + 2 * C};
  uint32_t* pn[2] = {gid + 3 * C, gid + 4 * C};
  uint32_t* scl = gid + 5 * C;
  uint32_t* scr = scl + C + 1;
  uint32_t* nst[2] = {scr + C + 1, scr + C + 1 + NC};
  uint32_t* ncn[2] = {nst[1] + NC, nst[1] + 2 * NC};
  uint32_t* nmode = ncn[1] + NC;
  uint32_t* nnl = nmode + NC;
  uint32_t* child = nnl + NC;
  uint32_t* piva = child + 2 * NC;
  uint32_t* pivb = piva + NC;
  const uint32_t l = (uint32_t)lane_id();
  const uint32_t b = block_id();
  if (b >= nsmall) return;  // block-uniform
  const uint32_t start = sm_start[b], c = sm_count[b], t = sm_tree[b], L0 = sm_level[b];
  const uint64_t ts = tseeds[t];
  uint32_t* tids = ids + (size_t)t * n + start;
  uint32_t* tl = lcnt_pos + (size_t)t * n + start;
  for (uint32_t e = l; e < c; e += WAVE) {
    gid[e] = tids[e];
    lid[0][e] = e;
    pn[0][e] = 0;
  }
  if (l == 0) {
    nst[0][0] = 0;
    ncn[0][0] = c;
  }
  __syncthreads();
  uint32_t A = 1;
  int cur = 0;
  for (uint32_t L = L0; A > 0; ++L) {
    const int nx = cur ^ 1;
    // pivots, one lane per node
    for (uint32_t v = l; v < A; v += WAVE) {
      const uint32_t s0 = nst[cur][v], cv = ncn[cur][v], gs0 = start + s0;
      const uint32_t* ln = lid[cur] + s0;
      const uint32_t ea = ln[tree_hash(ts, L, gs0, 0) % cv];
      uint32_t eb = ln[tree_hash(ts, L, gs0, 1) % cv];
      for (uint64_t r = 2; eb == ea && r < 10; ++r) eb = ln[tree_hash(ts, L, gs0, r) % cv];
      piva[v] = gid[ea];
      pivb[v] = gid[eb];
    }
    __syncthreads();
    // normals and the terms of the offset fold, one lane per (node, dimension)
    for (uint32_t x = l; x < A * ds; x += WAVE) {
      const uint32_t v = x / ds, j = x - v * ds;
      const float xa = Xs[(size_t)piva[v] * dps + j], xb = Xs[(size_t)pivb[v] * dps + j];
      const float nj = xa - xb;
      normals[(size_t)v * RSs + j] = nj;
      terms[(size_t)v * RSs + j] = nj * 0.5f * (xa + xb);
    }
    __syncthreads();
    // offsets: the sequential fold, one lane per node
    for (uint32_t v = l; v < A; v += WAVE) {
      float sacc = -0.0f;
      const float* tv = terms + (size_t)v * RSs;
      for (uint32_t j = 0; j < ds; ++j) sacc = sacc + tv[j];
      offs[v] = -sacc;
    }
    __syncthreads();
    // sides: the exact f32 margin (the subtree's rows are in cache; this phase
    // is bound by latency, not bytes)
    for (uint32_t pos = l; pos < c; pos += WAVE) {
      const uint32_t v = pn[cur][pos];
      uint32_t fl = 0, fr = 0;
      if (v != NONE) {
        const uint32_t g = gid[lid[cur][pos]];
        const float margin = dot_ref_row4(normals + (size_t)v * RSs, Xs + (size_t)g * dps, ds) + offs[v];
        if (fabsf(margin) < 1e-8f)
          fl = (tree_hash(ts, L, g, 2) & 1) == 0 ? 1u : 0u;
        else
          fl = margin > 0.0f ? 1u : 0u;
        fr = (tree_hash(ts, L, g, 3) & 1) == 0 ? 1u : 0u;
      }
      scl[pos] = fl;
      scr[pos] = fr;
    }
    __syncthreads();
    // exclusive scans over positions: lane l takes positions [4l, 4l + 4)
    {
      uint32_t vl[4], vr[4], sl_ = 0, sr_ = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t pos = 4 * l + u;
        vl[u] = pos < c ? scl[pos] : 0u;
        vr[u] = pos < c ? scr[pos] : 0u;
        sl_ += vl[u];
        sr_ += vr[u];
      }
      uint32_t xl = sl_, xr = sr_;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t yl = __shfl_up(xl, o, WAVE), yr = __shfl_up(xr, o, WAVE);
        if ((int)l >= o) {
          xl += yl;
          xr += yr;
        }
      }
      uint32_t el = xl - sl_, er = xr - sr_;
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t pos = 4 * l + u;
        if (pos < c) {
          scl[pos] = el;
          scr[pos] = er;
        }
        el += vl[u];
        er += vr[u];
      }
      if (l == WAVE - 1) {
        scl[c] = xl;
        scr[c] = xr;
      }
      if (l == 0) s_nA = 0;
    }
    __syncthreads();
    // split decisions, leaves and next-level nodes
    for (uint32_t v = l; v < A; v += WAVE) {
      const uint32_t s0 = nst[cur][v], cv = ncn[cur][v];
      const uint32_t cl = scl[s0 + cv] - scl[s0], rl = scr[s0 + cv] - scr[s0];
      uint32_t mode, nl;
      if (cl != 0 && cl != cv) {
        mode = 0;
        nl = cl;
      } else if (rl != 0 && rl != cv) {
        mode = 1;
        nl = rl;
      } else {
        mode = 2;
        nl = cv;
      }
      nmode[v] = mode;
      nnl[v] = nl;
      if (mode == 2) {
        tl[s0] = cv;
        child[2 * v] = NONE;
        child[2 * v + 1] = NONE;
        continue;
      }
      const bool deep = L + 1 > 100;
      const uint32_t cs[2] = {s0, s0 + nl};
      const uint32_t cc[2] = {nl, cv - nl};
      for (int s = 0; s < 2; ++s) {
        if (deep || cc[s] <= leaf_size) {
          tl[cs[s]] = cc[s];
          child[2 * v + s] = NONE;
        } else {
          const uint32_t x = atomicAdd(&s_nA, 1u);
          nst[nx][x] = cs[s];
          ncn[nx][x] = cc[s];
          child[2 * v + s] = x;
        }
      }
    }
    __syncthreads();
    // stable partition
    for (uint32_t pos = l; pos < c; pos += WAVE) {
      const uint32_t v = pn[cur][pos], e = lid[cur][pos];
      if (v == NONE || nmode[v] == 2) {
        lid[nx][pos] = e;
        pn[nx][pos] = NONE;
        continue;
      }
      const uint32_t* X = nmode[v] == 0 ? scl : scr;
      const uint32_t s0 = nst[cur][v];
      const uint32_t rank = X[pos] - X[s0];
      const bool left = (X[pos + 1] - X[pos]) != 0;
      const uint32_t np = left ? s0 + rank : s0 + nnl[v] + ((pos - s0) - rank);
      lid[nx][np] = e;
      pn[nx][np] = child[2 * v + (left ? 0 : 1)];
    }
    __syncthreads();
    A = s_nA;
    cur = nx;
    __syncthreads();
  }
  for (uint32_t pos = l; pos < c; pos += WAVE) tids[pos] = gid[lid[cur][pos]];
}

// Leaf list from the per-position counts: flags, then (after a scan of the
// flags into off) each leaf's start within its tree and its count. Leaves come
// out grouped by tree, in position order.
__global__ void k_leaf_flags(const uint32_t* __restrict__ lcnt_pos, size_t total,
                             uint32_t* __restrict__ flag) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) flag[e] = lcnt_pos[e] != 0 ? 1u : 0u;
}
__global__ void k_leaf_compact(const uint32_t* __restrict__ lcnt_pos, size_t total, uint32_t n,
                               const uint32_t* __restrict__ off, uint32_t* __restrict__ g_start,
                               uint32_t* __restrict__ g_count) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const uint32_t c = lcnt_pos[e];
  if (c == 0) return;
  const uint32_t o = off[e];
  g_start[o] = (uint32_t)(e % n);
  g_count[o] = c;
}
// Tree t's leaves start at off[t * n]; tb + 1 values gathered for the host.
__global__ void k_tree_offsets(const uint32_t* __restrict__ off, uint32_t n, uint32_t tb,
                               uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= tb) out[t] = off[(size_t)t * n];
}

// Leaf pairs of one tree. A group of LG lanes takes one leaf; a leaf of more
// than LG members is taken in consecutive chunks of LG (pairs within a chunk).
// The group stages its chunk's rows in LDS with coalesced 16-byte loads, at a
// row stride RS (floats, RS / 4 odd) that puts the group's rows on distinct
// banks. Each lane keeps its point's row (D = 50: in registers) and its graph
// row (KMAX keys, sorted) in registers, and inserts the keys of its
// chunk-mates: a key enters when it beats the row's worst and its (dist, idx)
// is not already present.
template <int D>
__device__ inline float reg_lds_sq_dist(const float (&me)[D > 0 ? D : 1], const float* other) {
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ch = 0; ch < D / 8; ++ch) {
    const float4 o0 = *reinterpret_cast<const float4*>(other + ch * 8);
    const float4 o1 = *reinterpret_cast<const float4*>(other + ch * 8 + 4);
    float t;
    t = me[ch * 8 + 0] - o0.x; acc[0] = acc[0] + t * t;
    t = me[ch * 8 + 1] - o0.y; acc[1] = acc[1] + t * t;
    t = me[ch * 8 + 2] - o0.z; acc[2] = acc[2] + t * t;
    t = me[ch * 8 + 3] - o0.w; acc[3] = acc[3] + t * t;
    t = me[ch * 8 + 4] - o1.x; acc[4] = acc[4] + t * t;
    t = me[ch * 8 + 5] - o1.y; acc[5] = acc[5] + t * t;
    t = me[ch * 8 + 6] - o1.z; acc[6] = acc[6] + t * t;
    t = me[ch * 8 + 7] - o1.w; acc[7] = acc[7] + t * t;
  }
  float tail = 0.f;
#pragma unroll
  for (int j = (D / 8) * 8; j < D; ++j) {
    const float t = me[j] - other[j];
    tail = tail + t * t;
  }
  return (((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]))) + tail;
}

// KE > 0: rows of exactly KE = KMAX entries (k == KE), so the row's worst is a
// fixed register and no test on k remains; KE = 0: rows of k <= KMAX entries.
// With KE, "already present" compares indices only: every key of neighbour q
// in the row carries the one distance computed for the pair, so an index match
// is a (dist, idx) match (an empty slot's index, 2^31 - 1, matches no point).
template <int D, int KMAX, int LG, int KE>
__global__ __launch_bounds__(256) void k_leaf_merge(
    const float* __restrict__ Xp, uint32_t d, uint32_t dp, uint32_t RS, uint32_t n,
    const uint32_t* __restrict__ ids_t, const uint32_t* __restrict__ g_start,
    const uint32_t* __restrict__ g_count, uint32_t nleaves, uint64_t* __restrict__ G, uint32_t k) {
  extern __shared__ __attribute__((aligned(16))) float s_rows[];  // [groups][LG][RS] | ids [groups][LG]
  __shared__ uint32_t s_rounds;
  const uint32_t GPB = blockDim.x / LG;
  const int g = threadIdx.x / LG, j = threadIdx.x % LG;
  const uint32_t li = block_id() * GPB + g;
  const bool lvalid = li < nleaves;
  const uint32_t start = lvalid ? g_start[li] : 0u, cnt = lvalid ? g_count[li] : 0u;
  const uint32_t nch = (cnt + LG - 1) / LG;
  if (threadIdx.x == 0) s_rounds = 0;
  __syncthreads();
  if (j == 0 && lvalid) atomicMax(&s_rounds, nch);
  __syncthreads();
  const uint32_t rounds = s_rounds;
  float* grows = s_rows + (size_t)g * LG * RS;
  uint32_t* gids = reinterpret_cast<uint32_t*>(s_rows + (size_t)GPB * LG * RS) + g * LG;
  const uint32_t c4 = (d + 3) / 4;
  constexpr int DR = D > 0 ? D : 1;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t c0 = r * LG;
    const uint32_t cc = (r < nch) ? min((uint32_t)LG, cnt - c0) : 0u;
    const bool valid = (uint32_t)j < cc;
    const uint32_t p = valid ? ids_t[start + c0 + j] : 0u;
    gids[j] = p;
    __syncthreads();
    for (uint32_t e = j; e < cc * c4; e += LG) {
      const uint32_t rr = e / c4, c = e - rr * c4;
      const float4 v = *reinterpret_cast<const float4*>(Xp + (size_t)gids[rr] * dp + 4 * c);
      *reinterpret_cast<float4*>(grows + (size_t)rr * RS + 4 * c) = v;
    }
    __syncthreads();
    if (valid) {
      static_assert(KE == 0 || KE == KMAX, "KE is the exact row length");
      const uint32_t kk = KE > 0 ? (uint32_t)KE : k;
      uint64_t row[KMAX];
      uint64_t* gr = G + (size_t)p * kk;
#pragma unroll
      for (int s = 0; s < KMAX; ++s) row[s] = (uint32_t)s < kk ? gr[s] : EMPTY_KEY;
      const float* mine = grows + (size_t)j * RS;
      float me[DR];
      if (D > 0) {
#pragma unroll
        for (int t = 0; t < DR / 4; ++t) {
          const float4 v = *reinterpret_cast<const float4*>(mine + 4 * t);
          me[4 * t] = v.x;
          me[4 * t + 1] = v.y;
          me[4 * t + 2] = v.z;
          me[4 * t + 3] = v.w;
        }
#pragma unroll
        for (int t = (DR / 4) * 4; t < DR; ++t) me[t] = mine[t];
      }
      for (uint32_t q = 0; q < cc; ++q) {
        if (q == (uint32_t)j) continue;
        const float* other = grows + (size_t)q * RS;
        const float dist = D > 0 ? reg_lds_sq_dist<D>(me, other) : sq_dist_ref(mine, other, (int)d);
        const uint32_t qi = gids[q];
        const uint64_t key = make_key(dist, qi, 1);
        uint64_t worst = EMPTY_KEY;
        if (KE > 0) {
          worst = row[KMAX - 1];
        } else {
#pragma unroll
          for (int s = 0; s < KMAX; ++s)
            if ((uint32_t)s == k - 1) worst = row[s];
        }
        if (key >= worst) continue;
        bool dup = false;
        if (KE > 0) {
          const uint32_t qk = qi << 1;
#pragma unroll
          for (int s = 0; s < KMAX; ++s) dup |= ((uint32_t)row[s] & ~1u) == qk;
        } else {
#pragma unroll
          for (int s = 0; s < KMAX; ++s) dup |= ((uint32_t)s < k) && ((row[s] >> 1) == (key >> 1));
        }
        if (dup) continue;
#pragma unroll
        for (int s = KMAX - 1; s > 0; --s) {
          if ((uint32_t)s < kk) {
            const uint64_t a = row[s - 1];
            row[s] = key < a ? a : (key < row[s] ? key : row[s]);
          }
        }
        row[0] = key < row[0] ? key : row[0];
      }
#pragma unroll
      for (int s = 0; s < KMAX; ++s)
        if ((uint32_t)s < kk) gr[s] = row[s];
    }
    __syncthreads();
  }
}

// Leaf pairs for rows of more than 64 entries (k <= 64 KV): one block of 256
// threads per leaf, the leaf taken in consecutive chunks of LG members as above.
// The chunk's rows are staged in LDS; wave w takes members j = w, w + 4, ...:
// lane q computes the key of chunk-mate q (q != j), and the row of j and those
// keys are sorted together in registers, equal (dist, idx) keep their first
// (the row's entry when present) and the row takes the first k. Same set
// function as the per-lane insertion of k_leaf_merge.
template <int D, int KV>
__global__ __launch_bounds__(256) void k_leaf_merge_wide(
    const float* __restrict__ Xp, uint32_t d, uint32_t dp, uint32_t RS, uint32_t n,
    const uint32_t* __restrict__ ids_t, const uint32_t* __restrict__ g_start,
    const uint32_t* __restrict__ g_count, uint32_t nleaves, uint64_t* __restrict__ G, uint32_t k,
    uint32_t LG) {
  constexpr int NT = KV * 2;  // row vectors + the candidates, rounded up to a power of two
  extern __shared__ __attribute__((aligned(16))) float s_rows[];  // [LG][RS] | ids[LG]
  const uint32_t li = block_id();
  if (li >= nleaves) return;  // block-uniform
  const uint32_t tid = threadIdx.x, l = (uint32_t)lane_id(), w = tid / WAVE;
  uint32_t* gids = reinterpret_cast<uint32_t*>(s_rows + (size_t)LG * RS);
  const uint32_t start = g_start[li], cnt = g_count[li];
  const uint32_t c4 = (d + 3) / 4;
  for (uint32_t c0 = 0; c0 < cnt; c0 += LG) {
    const uint32_t cc = min(LG, cnt - c0);
    __syncthreads();
    if (tid < cc) gids[tid] = ids_t[start + c0 + tid];
    __syncthreads();
    for (uint32_t e = tid; e < cc * c4; e += blockDim.x) {
      const uint32_t rr = e / c4, c = e - rr * c4;
      *reinterpret_cast<float4*>(s_rows + (size_t)rr * RS + 4 * c) =
          *reinterpret_cast<const float4*>(Xp + (size_t)gids[rr] * dp + 4 * c);
    }
    __syncthreads();
    for (uint32_t j = w; j < cc; j += blockDim.x / WAVE) {
      const float* mine = s_rows + (size_t)j * RS;
      uint64_t u[NT];
      uint64_t* gr = G + (size_t)gids[j] * k;
#pragma unroll
      for (int r = 0; r < KV; ++r) u[r] = (uint32_t)r * WAVE + l < k ? gr[r * WAVE + l] : EMPTY_KEY;
      u[KV] = EMPTY_KEY;
      if (l < cc && l != j) {
        const float dist = lds_sq_dist<D>(mine, s_rows + (size_t)l * RS, d);
        u[KV] = make_key(dist, gids[l], 1);
      }
#pragma unroll
      for (int r = KV + 1; r < NT; ++r) u[r] = EMPTY_KEY;
      wave_sort_n<NT>(u);
      uint32_t before = 0;
      uint64_t lastkey = EMPTY_KEY;
#pragma unroll
      for (int r = 0; r < NT; ++r) {
        uint64_t prev = shfl64(u[r], (int)((l + WAVE - 1) & (WAVE - 1)));
        if (l == 0) prev = lastkey;
        const bool keep = u[r] != EMPTY_KEY && !((r > 0 || l > 0) && (u[r] >> 1) == (prev >> 1));
        const uint64_t mask = __ballot(keep);
        const uint32_t rank = before + (uint32_t)__popcll(mask & ((1ull << l) - 1ull));
        if (keep && rank < k) gr[rank] = u[r];
        before += (uint32_t)__popcll(mask);
        lastkey = shfl64(u[r], WAVE - 1);
      }
      for (uint32_t pos = (before < k ? before : k) + l; pos < k; pos += WAVE) gr[pos] = EMPTY_KEY;
    }
  }
}
template __global__ void k_leaf_merge_wide<0, 2>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                                 const uint32_t*, const uint32_t*, const uint32_t*,
                                                 uint32_t, uint64_t*, uint32_t, uint32_t);
template __global__ void k_leaf_merge_wide<0, 4>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                                 const uint32_t*, const uint32_t*, const uint32_t*,
                                                 uint32_t, uint64_t*, uint32_t, uint32_t);
template __global__ void k_leaf_merge_wide<50, 2>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                                  const uint32_t*, const uint32_t*, const uint32_t*,
                                                  uint32_t, uint64_t*, uint32_t, uint32_t);
template __global__ void k_leaf_merge_wide<50, 4>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                                  const uint32_t*, const uint32_t*, const uint32_t*,
                                                  uint32_t, uint64_t*, uint32_t, uint32_t);

#define PSQ_LEAF_INST(D, KM, LG, KE)                                                              \
  template __global__ void k_leaf_merge<D, KM, LG, KE>(const float*, uint32_t, uint32_t, uint32_t, \
                                                       uint32_t, const uint32_t*, const uint32_t*, \
                                                       const uint32_t*, uint32_t, uint64_t*, uint32_t);
#define PSQ_LEAF_INST_D(D)     \
  PSQ_LEAF_INST(D, 16, 16, 0)  \
  PSQ_LEAF_INST(D, 32, 16, 0)  \
  PSQ_LEAF_INST(D, 32, 32, 0)  \
  PSQ_LEAF_INST(D, 64, 16, 0)  \
  PSQ_LEAF_INST(D, 64, 32, 0)  \
  PSQ_LEAF_INST(D, 64, 64, 0)  \
  PSQ_LEAF_INST(D, 15, 16, 15) \
  PSQ_LEAF_INST(D, 20, 16, 20) \
  PSQ_LEAF_INST(D, 26, 16, 26) \
  PSQ_LEAF_INST(D, 30, 16, 30)
PSQ_LEAF_INST_D(0)
PSQ_LEAF_INST_D(50)

}  // namespace psq
