// Oracle (test infrastructure only): the fuzzy connectivities of a kNN distance
// graph, restated from crates/polariseq-core/src/umap.rs.
//
//  * smooth_knn_dist (umap.rs:551-608): per row, rho = nearest non-zero
//    distance; sigma solves sum_j exp(-max(0, d_j - rho) / sigma) = log2(k+1)
//    by 64 bisection steps in f64 (tolerance 1e-5, doubling while the upper
//    bound is infinite), floored at 1e-3 of the mean distance (the mean counts
//    the self distance of zero) when rho > 0; weights exp(-(d - rho) / sigma)
//    as f32, 1 where d <= rho or sigma == 0.
//  * transpose by counting sort (umap.rs:357-381);
//  * per row, own entries and transposed entries, each sorted by column, merged
//    with the fuzzy union acc = acc + w - acc * w (own list first), entries with
//    acc <= 1e-5 dropped (umap.rs:488-537). A column can repeat within one list
//    only when an input row names a neighbour twice; the reference's order among
//    such entries is unspecified (sort_unstable_by_key), and here it is ascending
//    weight bits (sign masked), the order the GPU sort keys give.
// Compiled with -ffp-contract=off like the rest of the oracle.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

void smooth_knn_dist(const float* dists, size_t k, float* out) {
  if (k == 0) return;
  const size_t k_with_self = k + 1;
  const double target = std::log2((double)k_with_self);
  double rho = 0.0;
  for (size_t j = 0; j < k; ++j)
    if ((double)dists[j] > 0.0) {
      rho = (double)dists[j];
      break;
    }
  double lo = 0.0, hi = INFINITY, mid = 1.0;
  for (int it = 0; it < 64; ++it) {
    double psum = 0.0;
    for (size_t j = 0; j < k; ++j) {
      const double x = (double)dists[j] - rho;
      psum += x > 0.0 ? std::exp(-(x / mid)) : 1.0;
    }
    if (std::fabs(psum - target) < 1e-5) break;
    if (psum > target) {
      hi = mid;
      mid = (lo + hi) / 2.0;
    } else {
      lo = mid;
      mid = std::isinf(hi) ? mid * 2.0 : (lo + hi) / 2.0;
    }
  }
  double sum = 0.0;
  for (size_t j = 0; j < k; ++j) sum += (double)dists[j];
  const double mean = sum / (double)k_with_self;
  const double sigma = rho > 0.0 ? std::max(mid, 1e-3 * mean) : mid;
  for (size_t j = 0; j < k; ++j) {
    const double x = (double)dists[j] - rho;
    out[j] = (x <= 0.0 || sigma == 0.0) ? 1.0f : (float)std::exp(-(x / sigma));
  }
}

}  // namespace

extern "C" {

// Membership strengths of every row (aligned with the graph's entries).
void psq_o_fuzzy_strengths(const uint32_t* indptr, const float* dists, uint64_t n, float* out) {
  for (uint64_t i = 0; i < n; ++i)
    smooth_knn_dist(dists + indptr[i], indptr[i + 1] - indptr[i], out + indptr[i]);
}

// The symmetrised graph. Returns nnz, or -1 when cap is too small (nnz is still
// reported through *need).
int64_t psq_o_fuzzy_connectivities(const uint32_t* indptr, const uint32_t* indices,
                                   const float* dists, uint64_t n, uint32_t* out_indptr,
                                   uint32_t* out_indices, float* out_w, uint64_t cap,
                                   uint64_t* need) {
  const size_t nnz_in = indptr[n];
  std::vector<float> w(nnz_in);
  psq_o_fuzzy_strengths(indptr, dists, n, w.data());
  std::vector<uint32_t> t_indptr(n + 1, 0);
  for (size_t e = 0; e < nnz_in; ++e) t_indptr[indices[e] + 1]++;
  for (uint64_t j = 0; j < n; ++j) t_indptr[j + 1] += t_indptr[j];
  std::vector<uint32_t> t_cols(nnz_in);
  std::vector<float> t_vals(nnz_in);
  {
    std::vector<uint32_t> cur(t_indptr.begin(), t_indptr.end() - 1);
    for (uint64_t i = 0; i < n; ++i)
      for (uint32_t e = indptr[i]; e < indptr[i + 1]; ++e) {
        const uint32_t c = cur[indices[e]]++;
        t_cols[c] = (uint32_t)i;
        t_vals[c] = w[e];
      }
  }
  std::vector<std::vector<std::pair<uint32_t, float>>> rows(n);
  size_t total = 0;
  std::vector<std::pair<uint32_t, float>> own, tr;
  for (uint64_t i = 0; i < n; ++i) {
    own.clear();
    tr.clear();
    for (uint32_t e = indptr[i]; e < indptr[i + 1]; ++e) own.emplace_back(indices[e], w[e]);
    for (uint32_t e = t_indptr[i]; e < t_indptr[i + 1]; ++e) tr.emplace_back(t_cols[e], t_vals[e]);
    auto by_col = [](const std::pair<uint32_t, float>& a, const std::pair<uint32_t, float>& b) {
      if (a.first != b.first) return a.first < b.first;
      return (bits(a.second) & 0x7fffffffu) < (bits(b.second) & 0x7fffffffu);
    };
    std::sort(own.begin(), own.end(), by_col);
    std::sort(tr.begin(), tr.end(), by_col);
    size_t a = 0, b = 0;
    auto& out = rows[i];
    while (a < own.size() || b < tr.size()) {
      const uint32_t col =
          (a < own.size() && (b >= tr.size() || own[a].first <= tr[b].first)) ? own[a].first
                                                                              : tr[b].first;
      float acc = 0.0f;
      while (a < own.size() && own[a].first == col) {
        acc = acc + own[a].second - acc * own[a].second;
        ++a;
      }
      while (b < tr.size() && tr[b].first == col) {
        acc = acc + tr[b].second - acc * tr[b].second;
        ++b;
      }
      if (acc > 1e-5f) out.emplace_back(col, acc);
    }
    total += out.size();
  }
  if (need) *need = total;
  if (total > cap) return -1;
  out_indptr[0] = 0;
  size_t pos = 0;
  for (uint64_t i = 0; i < n; ++i) {
    for (auto& p : rows[i]) {
      out_indices[pos] = p.first;
      out_w[pos] = p.second;
      ++pos;
    }
    out_indptr[i + 1] = (uint32_t)pos;
  }
  return (int64_t)total;
}

}  // extern "C"
