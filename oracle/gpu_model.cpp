// Oracle (test infrastructure only): the GPU NN-Descent's semantics on the CPU.
//
// The GPU path reorganises the reference's algorithm (crates/polariseq-core/
// src/nndescent.rs:699-1003) so that every per-row step is a set function of
// its inputs, independent of thread order. This file states those semantics
// step by step. The HIP path's graph must equal this graph bit for bit for the
// same (data, options), which is what the GPU parity tests check. Its recall is
// judged separately against the reference restatement (nndescent_port.cpp).
//
// Differences from the reference, all deliberate:
//  * A row is kept sorted by the 64-bit key (dist_bits << 32 | idx << 1 | new).
//    An update is top-k of (row ∪ proposals), deduplicated on (dist, idx). The
//    reference's heap keeps the same k smallest; here exact ties go to the lower
//    index, not to the first arrival (nndescent.rs:176-181).
//  * Candidate sampling keeps, per row and flag class, the m smallest
//    (priority, neighbour) keys after deduplicating on neighbour. Of the two
//    copies of a neighbour j of row i in one class (i's own entry, and j's entry
//    naming i), the one from the earlier source row survives (the forward copy
//    when i < j): the reference applies the candidate updates in source-row
//    order and rejects a duplicate of a neighbour already held
//    (nndescent.rs:244-249, 856-892), so there the first arrival keeps its
//    priority (unless it was already evicted). Priorities are the same
//    hash_unit numerators (nndescent.rs:872), keyed by the row's sorted slot.
//  * Join proposals for a row are routed by a hash of the neighbour to one of
//    S/8 buckets of 8 slots. The row's current neighbours occupy ("pin") the
//    slots of their buckets (at most 8 per bucket), and a proposal of a
//    neighbour the row already holds is dropped. A bucket keeps the smallest
//    distinct proposal keys routed to it in one iteration, up to its free
//    slots, which is order-independent and so deterministic. A key of the
//    final top-k can be lost only when 8 smaller keys or neighbours share its
//    bucket. The reference buffers every proposal (nndescent.rs:925-968) and
//    drops duplicates when it pushes them (nndescent.rs:176-181).
//  * RP trees are built level by level. A node's hyperplane passes through two
//    members picked by hash from its segment; margins use the reference's
//    8-lane dot and the same offset fold (nndescent.rs:367-385, 439). A
//    degenerate split falls back to hash bits, not a Fisher-Yates shuffle. The
//    partition is stable.
//  * "changed" counts entries of the new top-k that were not in the row before.
#include "common.hpp"

#include <cmath>
#include <memory>

using namespace psq_oracle;

namespace {

constexpr uint32_t EMPTY = 0xFFFFFFFFu;
constexpr uint64_t EMPTY_KEY = ~0ull;
constexpr uint64_t GOLD = 0x9E3779B97F4A7C15ull;
constexpr size_t BUCKET = 8;  // slots per proposal bucket

static inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
static inline float bitsf(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static inline uint64_t make_key(float d, uint32_t idx, uint32_t is_new) {
  return ((uint64_t)fbits(d) << 32) | ((uint64_t)idx << 1) | is_new;
}
static inline uint32_t key_idx(uint64_t key) { return (uint32_t)(key & 0xFFFFFFFFull) >> 1; }
static inline float key_dist(uint64_t key) { return bitsf((uint32_t)(key >> 32)); }

// Hashes of the GPU forest and slots. Each is a splitmix chain (nndescent.rs:80-87).
static inline uint64_t tree_hash(uint64_t tseed, uint64_t level, uint64_t a, uint64_t salt) {
  return splitmix(tseed ^ splitmix(level * GOLD + salt) ^ splitmix(a * GOLD + 0x51ED27ull));
}
// Bucket of neighbour q among a row's n_buckets (a power of two, >= 8) in an
// iteration: the top bits of (q xor salt) * 0x9E3779B1, the salt a splitmix of
// (seed, iter).
static inline uint32_t slot_hash(uint64_t seed, uint64_t iter, uint32_t q, size_t n_buckets) {
  const uint32_t salt = (uint32_t)(splitmix((seed ^ 0xA5A5A5A5DEADBEEFull) + (iter << 32)) >> 32);
  uint32_t bits = 0;
  while (((size_t)1 << bits) < n_buckets) ++bits;
  return bits ? ((q ^ salt) * 0x9E3779B1u) >> (32 - bits) : 0u;
}

static inline uint32_t leaf_group(size_t leaf_size) {
  uint32_t g = 16;
  while (g < leaf_size && g < 64) g *= 2;
  return g;
}

struct Params {
  size_t n, d, k, m, n_iters, n_trees, leaf_size, slots, stop_at;
  uint64_t seed;
};

// Merge `props` (any order, any duplicates) into the sorted row `row[k]`.
// Returns how many entries of the new row were not in the old one.
static size_t merge_row(uint64_t* row, size_t k, std::vector<uint64_t>& props) {
  std::vector<uint64_t> all(row, row + k);
  all.insert(all.end(), props.begin(), props.end());
  std::sort(all.begin(), all.end());
  std::vector<uint64_t> out;
  out.reserve(k);
  uint64_t prev = EMPTY_KEY;
  for (uint64_t key : all) {
    if (key == EMPTY_KEY) break;
    if (!out.empty() && (key >> 1) == (prev >> 1)) continue;
    out.push_back(key);
    prev = key;
    if (out.size() == k) break;
  }
  size_t changed = 0;
  for (uint64_t key : out) {
    bool found = false;
    for (size_t s = 0; s < k; ++s)
      if ((row[s] >> 1) == (key >> 1) && row[s] != EMPTY_KEY) {
        found = true;
        break;
      }
    changed += !found;
  }
  for (size_t s = 0; s < k; ++s) row[s] = s < out.size() ? out[s] : EMPTY_KEY;
  return changed;
}

// One RP tree, level by level. Appends leaves as (start, count) into `ids`.
static void build_tree(const float* data, size_t d, const Params& P, uint64_t tseed,
                       std::vector<uint32_t>& ids, std::vector<std::pair<uint32_t, uint32_t>>& leaves) {
  const size_t n = P.n;
  ids.resize(n);
  for (size_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
  std::vector<std::pair<uint32_t, uint32_t>> level_nodes{{0u, (uint32_t)n}}, next;
  std::vector<float> normal(d);
  std::vector<uint32_t> tmp;
  std::vector<uint8_t> side, rside;
  for (uint64_t L = 0; !level_nodes.empty(); ++L) {
    next.clear();
    for (auto [start, c] : level_nodes) {
      if (c <= P.leaf_size || L > 100) {
        leaves.emplace_back(start, c);
        continue;
      }
      uint32_t* seg = ids.data() + start;
      // Pivots: two members picked by hash; up to 8 redraws for a distinct b.
      uint32_t a = seg[tree_hash(tseed, L, start, 0) % c];
      uint32_t b = seg[tree_hash(tseed, L, start, 1) % c];
      for (uint64_t r = 2; b == a && r < 10; ++r) b = seg[tree_hash(tseed, L, start, r) % c];
      const float* xa = data + (size_t)a * d;
      const float* xb = data + (size_t)b * d;
      for (size_t t = 0; t < d; ++t) normal[t] = xa[t] - xb[t];
      float s = -0.0f;
      for (size_t t = 0; t < d; ++t) s += normal[t] * 0.5f * (xa[t] + xb[t]);
      float offset = -s;
      side.assign(c, 0);
      rside.assign(c, 0);
      uint32_t cl = 0, rl = 0;
      for (uint32_t q = 0; q < c; ++q) {
        uint32_t i = seg[q];
        float margin = dot(normal.data(), data + (size_t)i * d, d) + offset;
        bool left;
        if (std::fabs(margin) < 1e-8f)
          left = (tree_hash(tseed, L, i, 2) & 1) == 0;
        else
          left = margin > 0.0f;
        side[q] = left;
        cl += left;
        rside[q] = (tree_hash(tseed, L, i, 3) & 1) == 0;
        rl += rside[q];
      }
      const std::vector<uint8_t>* use = &side;
      uint32_t nl = cl;
      if (cl == 0 || cl == c) {
        use = &rside;
        nl = rl;
        if (rl == 0 || rl == c) {
          leaves.emplace_back(start, c);
          continue;
        }
      }
      tmp.resize(c);
      uint32_t li = 0, ri = nl;
      for (uint32_t q = 0; q < c; ++q) {
        if ((*use)[q])
          tmp[li++] = seg[q];
        else
          tmp[ri++] = seg[q];
      }
      std::memcpy(seg, tmp.data(), c * sizeof(uint32_t));
      next.emplace_back(start, nl);
      next.emplace_back(start + nl, c - nl);
    }
    level_nodes.swap(next);
  }
}

}  // namespace

extern "C" {

struct psq_o_model_stats {
  uint64_t iterations;
  uint64_t n_trees, max_candidates, leaf_size, max_iterations, stop_at;
  uint64_t slots;
  uint64_t changed[64];
  uint64_t pairs[64];
  uint64_t proposals[64];
};

// init: low byte 0 = RP forest, 1 = random fill only, 2 = seed rows from
// init_graph (u32[n][cols], cols in bits 8..31 of init), 3 = RP forest whose
// hyperplanes split the f32 rows passed in init_graph (n x ds, ds in bits
// 8..31), e.g. learned codes, with leaf pairs measured on `data`.
int64_t psq_o_gpu_model_nndescent(const float* data, uint64_t n_obs, uint64_t n_dim, uint64_t k_req,
                                  uint64_t max_candidates, uint64_t max_iterations, double delta,
                                  uint64_t n_trees_req, uint64_t leaf_size_req, uint64_t slots_req,
                                  uint64_t heap_req, uint64_t seed, uint64_t init,
                                  const uint32_t* init_graph,
                                  int threads, uint32_t* indptr_out, uint32_t* indices_out,
                                  float* dists_out, psq_o_model_stats* stats) {
  if (n_obs == 0 || n_dim == 0 || k_req == 0) return -1;
  if (n_obs == 1) {
    // nndescent.rs:727 makes k = 1; the random fill's only draw is the point
    // itself (hash_index % 1 = 0, and (0 + 1) % 1 = 0, nndescent.rs:836-839),
    // pushed with its own distance into the empty row; nothing else changes it.
    indptr_out[0] = 0;
    indptr_out[1] = 1;
    indices_out[0] = 0;
    dists_out[0] = std::sqrt(sq_dist(data, data, n_dim));
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return 1;
  }
  Params P;
  P.n = n_obs;
  P.d = n_dim;
  // ko = neighbours output per row; k = entries each row keeps while iterating
  // (heap_req, at least ko; 0 -> ko). Options derived from k in the reference
  // (candidates, leaf size) follow ko; bucket slots and the stop rule follow k.
  const size_t ko = std::max<size_t>(1, std::min<size_t>(k_req, P.n - 1));
  P.k = std::min<size_t>(std::max<size_t>(ko, heap_req), P.n - 1);
  P.m = std::max<size_t>(1, max_candidates > 0 ? max_candidates : std::min<size_t>(2 * ko, 60));
  P.n_iters = max_iterations > 0 ? max_iterations
                                 : std::max<size_t>(5, (size_t)std::llround(std::log2((double)P.n)));
  P.n_trees = n_trees_req > 0
                  ? n_trees_req
                  : std::min<size_t>(48, 8 + (size_t)std::llround(std::sqrt((double)P.n) / 10.0));
  P.leaf_size = leaf_size_req > 0 ? leaf_size_req : std::max<size_t>(ko, 10);
  size_t S = 64;
  while (S < 4 * P.k) S *= 2;
  P.slots = slots_req > 0 ? slots_req : S;
  P.stop_at = (size_t)(delta * (double)P.n * (double)P.k);
  P.seed = seed;
  const size_t n = P.n, d = P.d, k = P.k, m = P.m;
  if ((P.slots & (P.slots - 1)) != 0 || P.slots < BUCKET) return -1;
  const int init_kind = (int)(init & 0xFF);
  const size_t init_cols = (size_t)(init >> 8);
  Pool pool(threads);
  auto row = [&](size_t i) { return data + i * d; };

  std::vector<uint64_t> G(n * k, EMPTY_KEY);

  // ── 1. initial rows ──
  if (init_kind == 0 || init_kind == 3) {
    if (init_kind == 3 && (!init_graph || init_cols == 0)) return -1;
    const float* sdata = init_kind == 3 ? reinterpret_cast<const float*>(init_graph) : data;
    const size_t sd = init_kind == 3 ? init_cols : d;
    std::vector<std::vector<uint32_t>> ids(P.n_trees);
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> leaves(P.n_trees);
    pool.parallel_for(P.n_trees, 1, [&](size_t t) {
      build_tree(sdata, sd, P, splitmix(seed ^ (uint64_t)(t + 1)), ids[t], leaves[t]);
    });
    // Leaf pairs, tree after tree. A leaf larger than the group size LG
    // (smallest power of two >= leaf_size, within [16, 64]) contributes pairs
    // inside consecutive chunks of LG members only.
    const uint32_t LG = leaf_group(P.leaf_size);
    for (size_t t = 0; t < P.n_trees; ++t) {
      pool.parallel_for(leaves[t].size(), 64, [&](size_t li) {
        auto [start, cnt] = leaves[t][li];
        std::vector<uint64_t> props;
        for (uint32_t c0 = 0; c0 < cnt; c0 += LG) {
          const uint32_t c = std::min(LG, cnt - c0);
          const uint32_t* mem = ids[t].data() + start + c0;
          for (uint32_t a = 0; a < c; ++a) {
            props.clear();
            uint32_t p = mem[a];
            for (uint32_t b = 0; b < c; ++b)
              if (b != a) props.push_back(make_key(sq_dist(row(p), row(mem[b]), d), mem[b], 1));
            merge_row(&G[(size_t)p * k], k, props);
          }
        }
      });
    }
  } else if (init_kind == 2) {
    if (!init_graph || init_cols == 0) return -1;
    pool.parallel_for(n, 256, [&](size_t p) {
      std::vector<uint64_t> props;
      for (size_t c = 0; c < init_cols; ++c) {
        uint32_t q = init_graph[p * init_cols + c];
        if (q == EMPTY || q == p || q >= n) continue;
        props.push_back(make_key(sq_dist(row(p), row(q), d), q, 1));
      }
      merge_row(&G[p * k], k, props);
    });
  }

  // Random fill of rows that still hold an empty slot (nndescent.rs:819-845):
  // attempts in order, each accepted unless its index is already present.
  pool.parallel_for(n, 256, [&](size_t i) {
    uint64_t* r = &G[i * k];
    size_t cnt = 0;
    while (cnt < k && r[cnt] != EMPTY_KEY) ++cnt;
    if (cnt == k) return;
    for (uint64_t attempt = 0; cnt < k && attempt < 64; ++attempt) {
      size_t j = hash_index(seed, i, attempt, n);
      if (j == i) j = (j + 1) % n;
      bool dup = false;
      for (size_t s = 0; s < cnt; ++s)
        if (key_idx(r[s]) == j) {
          dup = true;
          break;
        }
      if (dup) continue;
      float dd = sq_dist(row(i), row(j), d);
      if (dd >= INFINITY) continue;
      r[cnt++] = make_key(dd, (uint32_t)j, 1);
    }
    std::sort(r, r + k);
  });

  // ── 2. iterations ──
  std::vector<uint32_t> newl(n * m), oldl(n * m);
  std::vector<uint32_t> rcount(n + 1), roff(n + 1);
  std::vector<uint64_t> rev;  // (class << 63) | (prio24 << 32) | src
  std::vector<float> thr(n);
  const size_t n_buckets = P.slots / BUCKET;
  std::vector<std::vector<std::pair<uint32_t, uint64_t>>> emitted;  // per cell chunk
  size_t iters_done = 0;
  std::vector<uint64_t> log_changed, log_pairs, log_props;

  for (size_t iter = 0; iter < P.n_iters; ++iter) {
    // 2a. reverse proposals grouped by destination (order within a group is irrelevant)
    std::fill(rcount.begin(), rcount.end(), 0);
    for (size_t e = 0; e < n * k; ++e)
      if (G[e] != EMPTY_KEY) rcount[key_idx(G[e])]++;
    roff[0] = 0;
    for (size_t i = 0; i < n; ++i) roff[i + 1] = roff[i] + rcount[i];
    rev.assign(roff[n], 0);
    {
      std::vector<uint32_t> cur(roff.begin(), roff.end() - 1);
      for (size_t i = 0; i < n; ++i)
        for (size_t s = 0; s < k; ++s) {
          uint64_t key = G[i * k + s];
          if (key == EMPTY_KEY) continue;
          uint64_t cls = (key & 1) ? 0 : 1;  // new sorts first
          uint64_t pr = hash_prio24(seed, iter, i * k + s);
          rev[cur[key_idx(key)]++] = (cls << 63) | (pr << 32) | (uint64_t)i;
        }
    }
    // 2b. per row: dedup on neighbour per class (the copy from the earlier source
    // row wins), keep m smallest.
    pool.parallel_for(n, 256, [&](size_t i) {
      // (key, from_reverse)
      std::vector<std::pair<uint64_t, int>> all;
      for (size_t s = 0; s < k; ++s) {
        uint64_t key = G[i * k + s];
        if (key == EMPTY_KEY) continue;
        uint64_t cls = (key & 1) ? 0 : 1;
        uint64_t pr = hash_prio24(seed, iter, i * k + s);
        all.emplace_back((cls << 63) | (pr << 32) | key_idx(key), 0);
      }
      for (uint32_t r = roff[i]; r < roff[i + 1]; ++r) all.emplace_back(rev[r], 1);
      // group by (class, nbr); within a group the surviving copy first
      std::sort(all.begin(), all.end(), [i](const std::pair<uint64_t, int>& x,
                                           const std::pair<uint64_t, int>& y) {
        const uint64_t a = x.first, b = y.first;
        uint64_t ca = a >> 63, cb = b >> 63;
        uint32_t na = (uint32_t)a, nb = (uint32_t)b;
        if (ca != cb) return ca < cb;
        if (na != nb) return na < nb;
        // forward (source i) before reverse (source na) when i < na
        const bool fwd_first = i < na;
        return fwd_first ? x.second < y.second : x.second > y.second;
      });
      std::vector<uint64_t> uniq;
      for (size_t t = 0; t < all.size(); ++t) {
        if (t > 0 && (all[t].first >> 63) == (all[t - 1].first >> 63) &&
            (uint32_t)all[t].first == (uint32_t)all[t - 1].first)
          continue;
        uniq.push_back(all[t].first);
      }
      std::sort(uniq.begin(), uniq.end());
      size_t nn = 0, no = 0;
      for (uint64_t v : uniq) {
        if ((v >> 63) == 0) {
          if (nn < m) newl[i * m + nn++] = (uint32_t)v;
        } else {
          if (no < m) oldl[i * m + no++] = (uint32_t)v;
        }
      }
      for (; nn < m; ++nn) newl[i * m + nn] = EMPTY;
      for (; no < m; ++no) oldl[i * m + no] = EMPTY;
    });
    // 2c. flag update
    pool.parallel_for(n, 256, [&](size_t i) {
      for (size_t s = 0; s < k; ++s) {
        uint64_t& key = G[i * k + s];
        if (key == EMPTY_KEY || !(key & 1)) continue;
        uint32_t j = key_idx(key);
        for (size_t r = 0; r < m; ++r)
          if (newl[i * m + r] == j) {
            key &= ~1ull;
            break;
          }
      }
    });
    // 2d. thresholds and join into slots
    for (size_t i = 0; i < n; ++i) {
      uint64_t w = G[i * k + k - 1];
      thr[i] = w == EMPTY_KEY ? INFINITY : key_dist(w);
    }
    std::atomic<uint64_t> pairs{0}, props{0};
    const size_t CH = 256, nch = (n + CH - 1) / CH;
    emitted.assign(nch, {});
    pool.parallel_for(nch, 1, [&](size_t ci) {
      auto& out = emitted[ci];
      uint64_t lp = 0, lq = 0;
      auto emit = [&](uint32_t p, uint32_t q, float dd) {
        if (dd < thr[p]) { out.emplace_back(p, make_key(dd, q, 1)); ++lq; }
        if (dd < thr[q]) { out.emplace_back(q, make_key(dd, p, 1)); ++lq; }
      };
      for (size_t i = ci * CH; i < std::min(n, ci * CH + CH); ++i) {
        const uint32_t* N = &newl[i * m];
        const uint32_t* O = &oldl[i * m];
        for (size_t a = 0; a < m && N[a] != EMPTY; ++a) {
          uint32_t p = N[a];
          for (size_t b = a + 1; b < m && N[b] != EMPTY; ++b) {
            emit(p, N[b], sq_dist(row(p), row(N[b]), d));
            ++lp;
          }
          for (size_t b = 0; b < m && O[b] != EMPTY; ++b) {
            if (O[b] == p) continue;
            emit(p, O[b], sq_dist(row(p), row(O[b]), d));
            ++lp;
          }
        }
      }
      pairs += lp;
      props += lq;
    });
    // 2e. group proposals by destination; per bucket keep the BUCKET smallest
    //     distinct keys; merge them into the row.
    std::vector<uint32_t> pcount(n + 1, 0), poff(n + 1, 0);
    for (auto& v : emitted)
      for (auto& e : v) pcount[e.first]++;
    for (size_t i = 0; i < n; ++i) poff[i + 1] = poff[i] + pcount[i];
    std::vector<uint64_t> pk(poff[n]);
    {
      std::vector<uint32_t> cur(poff.begin(), poff.end() - 1);
      for (auto& v : emitted)
        for (auto& e : v) pk[cur[e.first]++] = e.second;
    }
    emitted.clear();
    std::vector<uint64_t> ch(n, 0);
    pool.parallel_for(n, 256, [&](size_t p) {
      if (poff[p] == poff[p + 1]) return;
      std::vector<std::vector<uint64_t>> bk(n_buckets);
      std::vector<size_t> pinned(n_buckets, 0);
      const uint64_t* rowp = &G[p * k];
      auto in_row = [&](uint32_t q) {
        for (size_t s = 0; s < k; ++s)
          if (rowp[s] != EMPTY_KEY && key_idx(rowp[s]) == q) return true;
        return false;
      };
      for (size_t s = 0; s < k; ++s)
        if (rowp[s] != EMPTY_KEY) pinned[slot_hash(seed, iter, key_idx(rowp[s]), n_buckets)]++;
      for (uint32_t r = poff[p]; r < poff[p + 1]; ++r) {
        uint64_t key = pk[r];
        if (in_row(key_idx(key))) continue;
        bk[slot_hash(seed, iter, key_idx(key), n_buckets)].push_back(key);
      }
      std::vector<uint64_t> pr;
      for (size_t bi = 0; bi < n_buckets; ++bi) {
        auto& b = bk[bi];
        const size_t cap = BUCKET - std::min(BUCKET, pinned[bi]);
        std::sort(b.begin(), b.end());
        b.erase(std::unique(b.begin(), b.end()), b.end());
        for (size_t t = 0; t < b.size() && t < cap; ++t) pr.push_back(b[t]);
      }
      ch[p] = merge_row(&G[p * k], k, pr);
    });
    size_t changed = 0;
    for (uint64_t v : ch) changed += v;
    ++iters_done;
    log_changed.push_back(changed);
    log_pairs.push_back(pairs.load());
    log_props.push_back(props.load());
    if (changed <= P.stop_at) break;
  }

  // ── 3. extraction: rows are sorted by (dist, idx) ──
  size_t pos = 0;
  indptr_out[0] = 0;
  for (size_t i = 0; i < n; ++i) {
    for (size_t s = 0; s < ko; ++s) {
      uint64_t key = G[i * k + s];
      if (key == EMPTY_KEY) break;
      indices_out[pos] = key_idx(key);
      dists_out[pos] = std::sqrt(key_dist(key));
      ++pos;
    }
    indptr_out[i + 1] = (uint32_t)pos;
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = iters_done;
    stats->n_trees = (init_kind == 0 || init_kind == 3) ? P.n_trees : 0;
    stats->max_candidates = m;
    stats->leaf_size = P.leaf_size;
    stats->max_iterations = P.n_iters;
    stats->stop_at = P.stop_at;
    stats->slots = P.slots;
    for (size_t i = 0; i < log_changed.size() && i < 64; ++i) {
      stats->changed[i] = log_changed[i];
      stats->pairs[i] = log_pairs[i];
      stats->proposals[i] = log_props[i];
    }
  }
  return (int64_t)pos;
}

}  // extern "C"
