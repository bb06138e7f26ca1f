// Oracle (test infrastructure only): exact kNN references and the scalar
// building blocks of the reference's distance and hashing code.
//
//  * psq_o_knn_exact64: FP64 brute force. Order = (distance, index), the
//    `by_distance` order of crates/polariseq-core/src/doublets.rs:544-546, with
//    the query itself excluded as in doublets.rs:557-651 (`j != i`). Distances
//    are Σ (double(a_t) − double(b_t))² summed sequentially over t with no
//    fused multiply-add, so any other FP64 implementation of the same sequence
//    (the GPU re-rank) gives the same bits. Pinned by the KATs of
//    doublets.rs:1702-1709 and neighbors.rs:248-265, by numpy float64
//    brute force, and by duplicate-point cases.
//  * psq_o_neighbors_exact: the reference's exact path, neighbors.rs:101-241
//    (bounded f32 max-heap per query, strict `<`, sort, sqrt, drop entry 0).
//  * scalar exports of sq_dist / dot (simd.rs:22-62) and the hashes
//    (nndescent.rs:80-100) so tests can pin them against published vectors.
#include "common.hpp"

#include <cmath>

using namespace psq_oracle;

extern "C" {

float psq_o_sq_dist(const float* a, const float* b, uint64_t d) { return sq_dist(a, b, d); }
float psq_o_dot(const float* a, const float* b, uint64_t d) { return dot(a, b, d); }
uint64_t psq_o_splitmix(uint64_t x) { return splitmix(x); }
float psq_o_hash_unit(uint64_t seed, uint64_t a, uint64_t b) { return hash_unit(seed, a, b); }
uint64_t psq_o_hash_index(uint64_t seed, uint64_t a, uint64_t b, uint64_t n) {
  return hash_index(seed, a, b, n);
}

double psq_o_sq_dist64(const float* a, const float* b, uint64_t d) {
  double s = 0.0;
  for (uint64_t t = 0; t < d; ++t) {
    double x = (double)a[t] - (double)b[t];
    s += x * x;
  }
  return s;
}

// For each query (an index into data, or 0..n_q-1 when queries is null): the k
// nearest other points by (d64, index). Rows with fewer than k other points are
// padded with index 0xFFFFFFFF and distance +inf.
int psq_o_knn_exact64(const float* data, uint64_t n, uint64_t d, const uint32_t* queries,
                      uint64_t n_q, uint64_t k, uint32_t* idx_out, double* d2_out, int threads) {
  if (n == 0 || d == 0 || k == 0) return -1;
  Pool pool(threads);
  pool.parallel_for(n_q, 16, [&](size_t qi) {
    uint64_t q = queries ? queries[qi] : qi;
    const float* xq = data + q * d;
    std::vector<std::pair<double, uint32_t>> all;
    all.reserve(n);
    for (uint64_t j = 0; j < n; ++j) {
      if (j == q) continue;
      all.emplace_back(psq_o_sq_dist64(xq, data + j * d, d), (uint32_t)j);
    }
    size_t kk = std::min<size_t>(k, all.size());
    std::partial_sort(all.begin(), all.begin() + kk, all.end());
    for (size_t s = 0; s < k; ++s) {
      if (s < kk) {
        idx_out[qi * k + s] = all[s].second;
        d2_out[qi * k + s] = all[s].first;
      } else {
        idx_out[qi * k + s] = 0xFFFFFFFFu;
        d2_out[qi * k + s] = INFINITY;
      }
    }
  });
  return 0;
}

// crates/polariseq-core/src/neighbors.rs:101-241. metric 0 = euclidean, 1 = cosine.
// Returns nnz, or -1 (EmptyInput), -2 (ZeroNeighbors).
int64_t psq_o_neighbors_exact(const float* emb, uint64_t n_obs, uint64_t n_dim,
                              uint64_t n_neighbors, int metric, uint32_t* indptr,
                              uint32_t* indices, float* dists, int threads) {
  if (n_obs == 0 || n_dim == 0) return -1;
  if (n_neighbors == 0) return -2;
  const size_t k = std::min<size_t>(n_neighbors, n_obs);
  struct Nb {
    float dist;
    uint32_t idx;
  };
  std::vector<Nb> per(n_obs * k);
  Pool pool(threads);
  pool.parallel_for(n_obs, 16, [&](size_t qi) {
    const float* qv = emb + qi * n_dim;
    std::vector<Nb> heap;
    heap.reserve(k);
    auto sift_down = [&](size_t p) {
      size_t len = heap.size();
      for (;;) {
        size_t l = 2 * p + 1, r = 2 * p + 2, m = p;
        if (l < len && heap[l].dist > heap[m].dist) m = l;
        if (r < len && heap[r].dist > heap[m].dist) m = r;
        if (m == p) break;
        std::swap(heap[p], heap[m]);
        p = m;
      }
    };
    for (size_t j = 0; j < n_obs; ++j) {
      const float* o = emb + j * n_dim;
      float dd;
      if (metric == 0) {
        dd = sq_dist(qv, o, n_dim);
      } else {
        float dt = dot(qv, o, n_dim), na = dot(qv, qv, n_dim), nb = dot(o, o, n_dim);
        float den = std::sqrt(na) * std::sqrt(nb);
        dd = den > 0.0f ? 1.0f - dt / den : 1.0f;
      }
      if (heap.size() < k) {
        heap.push_back({dd, (uint32_t)j});
        size_t c = heap.size() - 1;
        while (c > 0) {
          size_t par = (c - 1) / 2;
          if (heap[c].dist > heap[par].dist) {
            std::swap(heap[c], heap[par]);
            c = par;
          } else {
            break;
          }
        }
      } else if (dd < heap[0].dist) {
        heap[0] = {dd, (uint32_t)j};
        sift_down(0);
      }
    }
    // Rust's sort_by is stable; the heap order decides ties.
    std::stable_sort(heap.begin(), heap.end(),
                     [](const Nb& a, const Nb& b) { return a.dist < b.dist; });
    if (metric == 0)
      for (auto& nb : heap) nb.dist = std::sqrt(nb.dist);
    for (size_t s = 0; s < k; ++s) per[qi * k + s] = heap[s];
  });
  size_t pos = 0;
  indptr[0] = 0;
  for (size_t i = 0; i < n_obs; ++i) {
    for (size_t s = 1; s < k; ++s) {
      indices[pos] = per[i * k + s].idx;
      dists[pos] = per[i * k + s].dist;
      ++pos;
    }
    indptr[i + 1] = (uint32_t)pos;
  }
  return (int64_t)pos;
}

}  // extern "C"
