// Oracle (test infrastructure only): a step-by-step C++ restatement of the
// reference's CPU NN-Descent, crates/polariseq-core/src/nndescent.rs:41-1003.
//
// It serves two purposes and no other:
//   * the CPU baseline that bench.py times beside the GPU ("port" kind), and
//   * the recall yardstick: the GPU graph must reach at least its recall.
// The same hashing, auto-parameters, RP forest (with the gather copy), bucketed
// block-order apply, 8-lane sq_dist and stable row sort are followed, so the
// graph does not depend on the thread count (checked by a test).
//
// Parity status: the reference is Rust and cannot be compiled in this image, so
// this restatement is pinned by the reference's own invariant tests
// (nndescent.rs:1059-1199), by its recall thresholds, by the published
// SplitMix64 output vector and by the f64 checks of sq_dist (simd.rs:100-119).
// The exact bit-level graph of the Rust build is not pinned.
#include "common.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>

using namespace psq_oracle;

namespace {

constexpr uint32_t EMPTY = 0xFFFFFFFFu;
constexpr size_t N_BUCKETS = 64;   // nndescent.rs:275
constexpr size_t BLOCK = 4096;     // nndescent.rs:279
constexpr size_t GEN_CHUNK = 512;  // nndescent.rs:281
constexpr size_t GATHER_MAX = size_t(1) << 20;  // nndescent.rs:467

struct Update {  // nndescent.rs:286-290
  uint32_t dest, nbr;
  float dist;
};
struct CandUpdate {  // nndescent.rs:293-298
  uint32_t dest, nbr;
  float prio;
  bool is_new;
};

struct Rng {  // nndescent.rs:103-112
  uint64_t s;
  uint64_t next_u64() { s = splitmix(s); return s; }
  size_t below(size_t n) { return (size_t)(next_u64() % (uint64_t)n); }
};

// nndescent.rs:168-205: replace the root if better and not present, sift down.
static inline bool heap_push(uint32_t* idx, float* dist, uint8_t* flg, size_t k, uint32_t j,
                             float d, bool flag) {
  if (d >= dist[0]) return false;
  for (size_t s = 0; s < k; ++s)
    if (idx[s] == j) return false;
  idx[0] = j;
  dist[0] = d;
  flg[0] = flag;
  size_t p = 0;
  for (;;) {
    size_t l = 2 * p + 1, rr = 2 * p + 2, sw = p;
    if (l < k && dist[l] > dist[sw]) sw = l;
    if (rr < k && dist[rr] > dist[sw]) sw = rr;
    if (sw == p) break;
    std::swap(idx[p], idx[sw]);
    std::swap(dist[p], dist[sw]);
    std::swap(flg[p], flg[sw]);
    p = sw;
  }
  return true;
}

// nndescent.rs:244-269.
static inline void cand_push(uint32_t* idx, float* prio, size_t m, uint32_t j, float p) {
  if (p >= prio[0]) return;
  for (size_t s = 0; s < m; ++s)
    if (idx[s] == j) return;
  idx[0] = j;
  prio[0] = p;
  size_t q = 0;
  for (;;) {
    size_t l = 2 * q + 1, rr = 2 * q + 2, sw = q;
    if (l < m && prio[l] > prio[sw]) sw = l;
    if (rr < m && prio[rr] > prio[sw]) sw = rr;
    if (sw == q) break;
    std::swap(idx[q], idx[sw]);
    std::swap(prio[q], prio[sw]);
    q = sw;
  }
}

struct Ctx {
  const float* data;
  size_t n, d;
  size_t leaf_size, gather_max;
  const float* row(size_t i) const { return data + i * d; }
};

// nndescent.rs:367-385.
static void random_hyperplane(const float* data, size_t d, const uint32_t* ids, size_t n, Rng& rng,
                              std::vector<float>& normal, float& offset) {
  size_t a = ids[rng.below(n)];
  size_t b = ids[rng.below(n)];
  int tries = 0;
  while (b == a && tries < 8) {
    b = ids[rng.below(n)];
    ++tries;
  }
  const float* xa = data + a * d;
  const float* xb = data + b * d;
  normal.resize(d);
  for (size_t t = 0; t < d; ++t) normal[t] = xa[t] - xb[t];
  // A sequential left fold, as Rust's Iterator::sum over f32 (its neutral
  // element's sign cannot change the split: both signs give |margin| < 1e-8).
  float s = -0.0f;
  for (size_t t = 0; t < d; ++t) s += normal[t] * 0.5f * (xa[t] + xb[t]);
  offset = -s;
}

// nndescent.rs:389-405.
static size_t partition_in_place(uint32_t* ids, size_t n, const std::vector<uint8_t>& sides) {
  size_t lo = 0, hi = n;
  std::vector<uint32_t> pos(n);
  for (size_t i = 0; i < n; ++i) pos[i] = (uint32_t)i;
  while (lo < hi) {
    if (sides[pos[lo]]) {
      ++lo;
    } else {
      --hi;
      std::swap(ids[lo], ids[hi]);
      std::swap(pos[lo], pos[hi]);
    }
  }
  return lo;
}

// nndescent.rs:408-415.
static size_t random_split(uint32_t* ids, size_t n, Rng& rng) {
  for (size_t i = n; i-- > 1;) {
    size_t j = rng.below(i + 1);
    std::swap(ids[i], ids[j]);
  }
  return n / 2;
}

using Leaves = std::vector<std::pair<uint32_t, uint32_t>>;

// nndescent.rs:419-458.
static void rp_tree_leaves(const float* data, size_t d, uint32_t* ids, size_t n, size_t base,
                           size_t leaf_size, Rng& rng, size_t depth, Leaves& out) {
  if (n <= leaf_size || depth > 100) {
    out.emplace_back((uint32_t)base, (uint32_t)(base + n));
    return;
  }
  std::vector<float> normal;
  float offset;
  random_hyperplane(data, d, ids, n, rng, normal, offset);
  std::vector<uint8_t> sides(n);
  for (size_t q = 0; q < n; ++q) {
    float margin = dot(normal.data(), data + (size_t)ids[q] * d, d) + offset;
    if (std::fabs(margin) < 1e-8f)
      sides[q] = (rng.next_u64() & 1) == 0;
    else
      sides[q] = margin > 0.0f;
  }
  size_t split = partition_in_place(ids, n, sides);
  if (split == 0 || split == n) {
    split = random_split(ids, n, rng);
    if (split == 0 || split == n) {
      out.emplace_back((uint32_t)base, (uint32_t)(base + n));
      return;
    }
  }
  rp_tree_leaves(data, d, ids, split, base, leaf_size, rng, depth + 1, out);
  rp_tree_leaves(data, d, ids + split, n - split, base + split, leaf_size, rng, depth + 1, out);
}

// nndescent.rs:474-549. The copy below gather_max is a memory layout choice;
// the recursion draws from the same rng in the same order either way.
static void rp_tree_sequential(const Ctx& c, uint32_t* ids, size_t n, size_t base, Rng& rng,
                               size_t depth, Leaves& out) {
  const size_t d = c.d;
  if (n <= c.gather_max) {
    std::vector<float> local(n * d);
    for (size_t q = 0; q < n; ++q) std::memcpy(&local[q * d], c.row(ids[q]), d * sizeof(float));
    std::vector<uint32_t> local_ids(n);
    for (size_t q = 0; q < n; ++q) local_ids[q] = (uint32_t)q;
    rp_tree_leaves(local.data(), d, local_ids.data(), n, base, c.leaf_size, rng, depth, out);
    std::vector<uint32_t> snap(ids, ids + n);
    for (size_t q = 0; q < n; ++q) ids[q] = snap[local_ids[q]];
    return;
  }
  if (n <= c.leaf_size || depth > 100) {
    out.emplace_back((uint32_t)base, (uint32_t)(base + n));
    return;
  }
  std::vector<float> normal;
  float offset;
  random_hyperplane(c.data, d, ids, n, rng, normal, offset);
  std::vector<uint8_t> sides(n);
  for (size_t q = 0; q < n; ++q) {
    float margin = dot(normal.data(), c.row(ids[q]), d) + offset;
    if (std::fabs(margin) < 1e-8f)
      sides[q] = (rng.next_u64() & 1) == 0;
    else
      sides[q] = margin > 0.0f;
  }
  size_t split = partition_in_place(ids, n, sides);
  if (split == 0 || split == n) {
    split = random_split(ids, n, rng);
    if (split == 0 || split == n) {
      out.emplace_back((uint32_t)base, (uint32_t)(base + n));
      return;
    }
  }
  rp_tree_sequential(c, ids, split, base, rng, depth + 1, out);
  rp_tree_sequential(c, ids + split, n - split, base + split, rng, depth + 1, out);
}

// nndescent.rs:555-628. The two halves are independent (their own seeds), so
// running them one after the other gives the same leaves as rayon::join.
static void rp_tree_parallel(const Ctx& c, uint32_t* ids, size_t n, size_t base, uint64_t seed,
                             size_t depth, Leaves& out) {
  const size_t d = c.d;
  if (n < 8192 || depth >= 6) {
    Rng rng{seed};
    rp_tree_sequential(c, ids, n, base, rng, depth, out);
    return;
  }
  Rng rng{seed};
  std::vector<float> normal;
  float offset;
  random_hyperplane(c.data, d, ids, n, rng, normal, offset);
  std::vector<uint8_t> sides(n);
  for (size_t q = 0; q < n; ++q) {
    uint32_t i = ids[q];
    float margin = dot(normal.data(), c.row(i), d) + offset;
    if (std::fabs(margin) < 1e-8f)
      sides[q] = (splitmix(seed ^ (uint64_t)i) & 1) == 0;
    else
      sides[q] = margin > 0.0f;
  }
  size_t split = partition_in_place(ids, n, sides);
  if (split == 0 || split == n) split = random_split(ids, n, rng);
  rp_tree_parallel(c, ids, split, base, splitmix(seed ^ 1), depth + 1, out);
  rp_tree_parallel(c, ids + split, n - split, base + split, splitmix(seed ^ 2), depth + 1, out);
}

struct Tree {
  std::vector<uint32_t> ids;
  Leaves leaves;
};

}  // namespace

extern "C" {

// nndescent.rs:648-650.
uint64_t psq_o_effective_k(uint64_t n_obs, uint64_t k) {
  uint64_t nm1 = n_obs > 0 ? n_obs - 1 : 0;
  uint64_t e = k < nm1 ? k : nm1;
  return e > 0 ? e : 1;
}

// Per-run statistics of the last port run on this thread.
struct psq_o_port_stats {
  uint64_t iterations;
  uint64_t n_trees, max_candidates, leaf_size, max_iterations, stop_at;
  double seconds_forest, seconds_iterations, seconds_total;
  uint64_t changed[64];
};

// nndescent.rs:699-1003. Returns nnz; -1 on invalid input (the Rust code panics).
int64_t psq_o_nndescent_port(const float* data, uint64_t n_obs, uint64_t n_dim, uint64_t k_req,
                             uint64_t max_candidates, uint64_t max_iterations, double delta,
                             uint64_t n_trees_req, uint64_t leaf_size_req, uint64_t gather_max_req,
                             uint64_t seed, int threads, uint32_t* indptr_out,
                             uint32_t* indices_out, float* dists_out, psq_o_port_stats* stats) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  if (n_obs == 0 || n_dim == 0 || k_req == 0) return -1;
  const size_t n = n_obs, d = n_dim;
  const size_t k = std::max<size_t>(1, std::min<size_t>(k_req, n - 1));
  const size_t max_cand = std::max<size_t>(1, max_candidates > 0 ? max_candidates
                                                                   : std::min<size_t>(2 * k, 60));
  const size_t n_iters = max_iterations > 0
                             ? max_iterations
                             : std::max<size_t>(5, (size_t)std::llround(std::log2((double)n)));
  const size_t n_trees =
      n_trees_req > 0 ? n_trees_req
                      : std::min<size_t>(48, 8 + (size_t)std::llround(std::sqrt((double)n) / 10.0));
  const size_t leaf_size = leaf_size_req > 0 ? leaf_size_req : std::max<size_t>(k, 10);
  const size_t gather_max = gather_max_req > 0 ? gather_max_req : GATHER_MAX;
  const size_t rpb = std::max<size_t>(1, (n + N_BUCKETS - 1) / N_BUCKETS);
  Ctx c{data, n, d, leaf_size, gather_max};
  Pool pool(threads);

  std::vector<uint32_t> hidx(n * k, EMPTY);
  std::vector<float> hdist(n * k, INFINITY);
  std::vector<uint8_t> hflag(n * k, 0);

  // ── 1. forest (nndescent.rs:767-817) ──
  std::vector<Tree> forest(n_trees);
  pool.parallel_for(n_trees, 1, [&](size_t t) {
    Tree& tr = forest[t];
    tr.ids.resize(n);
    for (size_t i = 0; i < n; ++i) tr.ids[i] = (uint32_t)i;
    rp_tree_parallel(c, tr.ids.data(), n, 0, splitmix(seed ^ (uint64_t)(t + 1)), 0, tr.leaves);
  });

  // Bucketed apply, nndescent.rs:307-324: per bucket, chunks in order.
  auto apply_updates = [&](std::vector<std::vector<std::vector<Update>>>& gen) -> size_t {
    std::vector<size_t> changed(N_BUCKETS, 0);
    pool.parallel_for(N_BUCKETS, 1, [&](size_t b) {
      size_t ch = 0;
      for (auto& chunk : gen)
        for (const Update& u : chunk[b])
          if (heap_push(&hidx[(size_t)u.dest * k], &hdist[(size_t)u.dest * k],
                        &hflag[(size_t)u.dest * k], k, u.nbr, u.dist, true))
            ++ch;
      changed[b] = ch;
    });
    size_t s = 0;
    for (size_t v : changed) s += v;
    return s;
  };

  for (size_t t = 0; t < n_trees; ++t) {
    const Tree& tr = forest[t];
    size_t nchunks = (tr.leaves.size() + 63) / 64;
    std::vector<std::vector<std::vector<Update>>> gen(nchunks,
                                                      std::vector<std::vector<Update>>(N_BUCKETS));
    pool.parallel_for(nchunks, 1, [&](size_t ci) {
      auto& buckets = gen[ci];
      size_t l0 = ci * 64, l1 = std::min(tr.leaves.size(), l0 + 64);
      for (size_t li = l0; li < l1; ++li) {
        const uint32_t* leaf = tr.ids.data() + tr.leaves[li].first;
        size_t len = tr.leaves[li].second - tr.leaves[li].first;
        for (size_t a = 0; a < len; ++a) {
          uint32_t p = leaf[a];
          for (size_t b = a + 1; b < len; ++b) {
            uint32_t q = leaf[b];
            float dd = sq_dist(c.row(p), c.row(q), d);
            buckets[p / rpb].push_back({p, q, dd});
            buckets[q / rpb].push_back({q, p, dd});
          }
        }
      }
    });
    apply_updates(gen);
  }
  forest.clear();
  forest.shrink_to_fit();

  // Random fill of empty slots, nndescent.rs:819-845.
  pool.parallel_for(n, 256, [&](size_t i) {
    uint32_t* idx = &hidx[i * k];
    uint64_t attempt = 0;
    auto has_empty = [&] {
      for (size_t s = 0; s < k; ++s)
        if (idx[s] == EMPTY) return true;
      return false;
    };
    while (has_empty() && attempt < 64) {
      size_t j = hash_index(seed, i, attempt, n);
      if (j == i) j = (j + 1) % n;
      float dd = sq_dist(c.row(i), c.row(j), d);
      heap_push(idx, &hdist[i * k], &hflag[i * k], k, (uint32_t)j, dd, true);
      ++attempt;
    }
  });
  auto t_forest = clk::now();

  // ── 2. iterations (nndescent.rs:847-975) ──
  const size_t m = max_cand;
  std::vector<uint32_t> nc_idx(n * m), oc_idx(n * m);
  std::vector<float> nc_prio(n * m), oc_prio(n * m);
  const size_t stop_at = (size_t)(delta * (double)n * (double)k);
  std::vector<float> thr(n);
  size_t iters_done = 0;
  std::vector<uint64_t> changed_log;

  for (size_t iter = 0; iter < n_iters; ++iter) {
    std::fill(nc_idx.begin(), nc_idx.end(), EMPTY);
    std::fill(oc_idx.begin(), oc_idx.end(), EMPTY);
    std::fill(nc_prio.begin(), nc_prio.end(), INFINITY);
    std::fill(oc_prio.begin(), oc_prio.end(), INFINITY);
    for (size_t b0 = 0; b0 < n; b0 += BLOCK) {
      size_t b1 = std::min(n, b0 + BLOCK);
      size_t nchunks = (b1 - b0 + GEN_CHUNK - 1) / GEN_CHUNK;
      std::vector<std::vector<std::vector<CandUpdate>>> gen(
          nchunks, std::vector<std::vector<CandUpdate>>(N_BUCKETS));
      pool.parallel_for(nchunks, 1, [&](size_t ci) {
        auto& buckets = gen[ci];
        size_t c0 = b0 + ci * GEN_CHUNK, c1 = std::min(b1, c0 + GEN_CHUNK);
        for (size_t i = c0; i < c1; ++i) {
          for (size_t s = 0; s < k; ++s) {
            uint32_t j = hidx[i * k + s];
            if (j == EMPTY) continue;
            bool is_new = hflag[i * k + s] != 0;
            float prio = hash_unit(seed, iter, i * k + s);
            buckets[i / rpb].push_back({(uint32_t)i, j, prio, is_new});
            buckets[j / rpb].push_back({j, (uint32_t)i, prio, is_new});
          }
        }
      });
      // nndescent.rs:326-354, new then old.
      pool.parallel_for(N_BUCKETS, 1, [&](size_t b) {
        for (auto& chunk : gen)
          for (const CandUpdate& u : chunk[b]) {
            if (u.is_new)
              cand_push(&nc_idx[(size_t)u.dest * m], &nc_prio[(size_t)u.dest * m], m, u.nbr,
                        u.prio);
          }
        for (auto& chunk : gen)
          for (const CandUpdate& u : chunk[b]) {
            if (!u.is_new)
              cand_push(&oc_idx[(size_t)u.dest * m], &oc_prio[(size_t)u.dest * m], m, u.nbr,
                        u.prio);
          }
      });
    }
    // Flag update, nndescent.rs:894-909.
    pool.parallel_for(n, 1024, [&](size_t i) {
      const uint32_t* cands = &nc_idx[i * m];
      for (size_t s = 0; s < k; ++s) {
        uint32_t j = hidx[i * k + s];
        if (hflag[i * k + s] && j != EMPTY) {
          for (size_t r = 0; r < m; ++r)
            if (cands[r] == j) {
              hflag[i * k + s] = 0;
              break;
            }
        }
      }
    });
    // Local join, nndescent.rs:912-969.
    for (size_t i = 0; i < n; ++i) thr[i] = hdist[i * k];
    size_t changed = 0;
    for (size_t b0 = 0; b0 < n; b0 += BLOCK) {
      size_t b1 = std::min(n, b0 + BLOCK);
      size_t nchunks = (b1 - b0 + GEN_CHUNK - 1) / GEN_CHUNK;
      std::vector<std::vector<std::vector<Update>>> gen(nchunks,
                                                        std::vector<std::vector<Update>>(N_BUCKETS));
      pool.parallel_for(nchunks, 1, [&](size_t ci) {
        auto& buckets = gen[ci];
        auto emit = [&](uint32_t p, uint32_t q, float dd) {
          if (dd < thr[p]) buckets[p / rpb].push_back({p, q, dd});
          if (dd < thr[q]) buckets[q / rpb].push_back({q, p, dd});
        };
        size_t c0 = b0 + ci * GEN_CHUNK, c1 = std::min(b1, c0 + GEN_CHUNK);
        for (size_t i = c0; i < c1; ++i) {
          const uint32_t* news = &nc_idx[i * m];
          const uint32_t* olds = &oc_idx[i * m];
          for (size_t a = 0; a < m; ++a) {
            uint32_t p = news[a];
            if (p == EMPTY) continue;
            const float* xp = c.row(p);
            for (size_t b = a + 1; b < m; ++b) {
              uint32_t q = news[b];
              if (q == EMPTY || q == p) continue;
              emit(p, q, sq_dist(xp, c.row(q), d));
            }
            for (size_t b = 0; b < m; ++b) {
              uint32_t q = olds[b];
              if (q == EMPTY || q == p) continue;
              emit(p, q, sq_dist(xp, c.row(q), d));
            }
          }
        }
      });
      changed += apply_updates(gen);
    }
    ++iters_done;
    changed_log.push_back(changed);
    if (changed <= stop_at) break;
  }
  auto t_iter = clk::now();

  // ── 3. extraction (nndescent.rs:977-1002) ──
  indptr_out[0] = 0;
  size_t pos = 0;
  std::vector<std::pair<uint32_t, float>> rowbuf;
  rowbuf.reserve(k);
  for (size_t i = 0; i < n; ++i) {
    rowbuf.clear();
    for (size_t s = 0; s < k; ++s) {
      uint32_t j = hidx[i * k + s];
      if (j != EMPTY) rowbuf.emplace_back(j, hdist[i * k + s]);
    }
    std::stable_sort(rowbuf.begin(), rowbuf.end(),
                     [](const std::pair<uint32_t, float>& a, const std::pair<uint32_t, float>& b) {
                       return a.second < b.second;
                     });
    for (auto& e : rowbuf) {
      indices_out[pos] = e.first;
      dists_out[pos] = std::sqrt(e.second);
      ++pos;
    }
    indptr_out[i + 1] = (uint32_t)pos;
  }
  auto t_end = clk::now();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = iters_done;
    stats->n_trees = n_trees;
    stats->max_candidates = max_cand;
    stats->leaf_size = leaf_size;
    stats->max_iterations = n_iters;
    stats->stop_at = stop_at;
    stats->seconds_forest = std::chrono::duration<double>(t_forest - t0).count();
    stats->seconds_iterations = std::chrono::duration<double>(t_iter - t_forest).count();
    stats->seconds_total = std::chrono::duration<double>(t_end - t0).count();
    for (size_t i = 0; i < changed_log.size() && i < 64; ++i) stats->changed[i] = changed_log[i];
  }
  return (int64_t)pos;
}

}  // extern "C"
