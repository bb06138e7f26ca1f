// Oracle (test infrastructure only): shared CPU helpers.
//
// Nothing here is part of the product. Only tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg load the library built from these files, and only
// as a checker or as the timed CPU baseline.
//
// The arithmetic below restates the reference's Rust code; every function cites
// the file:line it follows (paths relative to the reference repository root).
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace psq_oracle {

// crates/polariseq-core/src/nndescent.rs:80-87 (splitmix64 finaliser of x + golden ratio).
static inline uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// crates/polariseq-core/src/nndescent.rs:91-95. The 24-bit numerator is exact in f32.
static inline uint32_t hash_prio24(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t h = splitmix(seed ^ splitmix(a) ^ splitmix(b * 0x9E3779B97F4A7C15ull));
  return (uint32_t)(h >> 40);
}
static inline float hash_unit(uint64_t seed, uint64_t a, uint64_t b) {
  return (float)hash_prio24(seed, a, b) / (float)(1u << 24);
}

// crates/polariseq-core/src/nndescent.rs:97-100.
static inline uint64_t hash_index(uint64_t seed, uint64_t a, uint64_t b, uint64_t n) {
  return splitmix(seed ^ splitmix(a) ^ splitmix(b + 0x12345678ull)) % n;
}

// crates/polariseq-core/src/simd.rs:79-81: fixed pairwise reduce of the 8 lanes.
static inline float reduce8(const float* acc) {
  return ((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]));
}

// crates/polariseq-core/src/simd.rs:45-62: 8 independent lane sums over chunks of
// 8, fixed reduce, then a sequential tail added last. Built with
// -ffp-contract=off, because Rust never fuses the multiply and the add.
static inline float sq_dist(const float* a, const float* b, size_t d) {
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t c = d / 8;
  for (size_t ch = 0; ch < c; ++ch) {
    for (int l = 0; l < 8; ++l) {
      float t = a[ch * 8 + l] - b[ch * 8 + l];
      acc[l] += t * t;
    }
  }
  float tail = 0.0f;
  for (size_t i = c * 8; i < d; ++i) {
    float t = a[i] - b[i];
    tail += t * t;
  }
  return reduce8(acc) + tail;
}

// crates/polariseq-core/src/simd.rs:22-37.
static inline float dot(const float* a, const float* b, size_t d) {
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t c = d / 8;
  for (size_t ch = 0; ch < c; ++ch) {
    for (int l = 0; l < 8; ++l) acc[l] += a[ch * 8 + l] * b[ch * 8 + l];
  }
  float tail = 0.0f;
  for (size_t i = c * 8; i < d; ++i) tail += a[i] * b[i];
  return reduce8(acc) + tail;
}

// A small persistent pool: parallel_for hands out indices [0, n) in chunks of
// `grain`. Results never depend on which worker runs which index, because every
// caller writes per-index outputs only.
class Pool {
 public:
  explicit Pool(int threads) : nthreads_(threads < 1 ? 1 : threads) {
    for (int t = 1; t < nthreads_; ++t) workers_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  int threads() const { return nthreads_; }

  void parallel_for(size_t n, size_t grain, const std::function<void(size_t)>& fn) {
    if (n == 0) return;
    if (nthreads_ == 1 || n <= grain) {
      for (size_t i = 0; i < n; ++i) fn(i);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = &fn;
      n_ = n;
      grain_ = grain;
      next_.store(0);
      active_ = nthreads_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return active_ == 0; });
    job_ = nullptr;
  }

 private:
  void work() {
    for (;;) {
      size_t s = next_.fetch_add(grain_);
      if (s >= n_) break;
      size_t e = std::min(n_, s + grain_);
      for (size_t i = s; i < e; ++i) (*job_)(i);
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      work();
      {
        std::lock_guard<std::mutex> g(mu_);
        if (--active_ == 0) done_cv_.notify_all();
      }
    }
  }
  int nthreads_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t)>* job_ = nullptr;
  size_t n_ = 0, grain_ = 1;
  std::atomic<size_t> next_{0};
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

}  // namespace psq_oracle
