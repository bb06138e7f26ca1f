// Kernel declarations shared by the kernel sources and the host orchestration.
#pragma once
#include "psq_device.hpp"

namespace psq {

constexpr int SCAN_BLOCK = 1024;  // threads per scan block

__global__ void k_fill_u64(uint64_t* p, size_t count, uint64_t v);
__global__ void k_fill_u32(uint32_t* p, size_t count, uint32_t v);
__global__ void k_fill_u8(uint8_t* p, size_t count, uint8_t v);
__global__ void k_pad_rows(const float* X, uint32_t n, uint32_t d, uint32_t dp, float* Xp);
__global__ void k_scan_local(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* block_sums);
__global__ void k_scan_sums(uint32_t* sums, uint32_t nb, uint32_t* total_out);
__global__ void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* sums);

__global__ void k_random_fill(const float* Xp, uint32_t n, uint32_t d, uint32_t dp, uint64_t* G,
                              uint32_t k, uint64_t seed, uint8_t* touched);
__global__ void k_rev_count(const uint64_t* G, size_t nk, uint32_t* rcount);
__global__ void k_rev_scatter(const uint64_t* G, uint32_t n, uint32_t k, uint64_t seed,
                              uint64_t iter, const uint32_t* roff, uint32_t* rcount, uint64_t* rev);
__global__ void k_sample_select(uint64_t* G, uint32_t n, uint32_t k, uint32_t m, uint64_t seed,
                                uint64_t iter, const uint32_t* roff, const uint64_t* rev,
                                uint32_t* newl, uint32_t* oldl, uint32_t* ncnt);
template <int D, int CPL>
__global__ void k_join(const float* Xp, uint32_t n, uint32_t d, uint32_t dp, const uint32_t* newl,
                       const uint32_t* oldl, const uint32_t* ncnt, uint32_t m, const float* thr,
                       uint64_t* slots, uint8_t* touched, uint32_t S, uint64_t seed, uint64_t iter,
                       unsigned long long* counters);
template <int NR>
__global__ void k_merge(uint64_t* G, uint32_t n, uint32_t k, uint64_t* slots, uint32_t S,
                        uint8_t* touched, float* thr, int force, unsigned long long* changed);
__global__ void k_row_counts(const uint64_t* G, uint32_t n, uint32_t k, uint32_t* cnt);
__global__ void k_extract(const uint64_t* G, uint32_t n, uint32_t k, const uint32_t* off,
                          uint32_t* indices, float* dists);

// exact kNN (exact_kernels.hip)
template <int D, int K>
__global__ void k_exact_pass1(const float* Xp, uint32_t n, uint32_t d, uint32_t dp,
                              const uint32_t* queries, uint32_t nq, uint32_t seg_len,
                              uint32_t nseg, float* partial);
__global__ void k_exact_tau(const float* partial, uint32_t nq, uint32_t nseg, uint32_t K,
                            uint32_t k, uint32_t d, float* T);
template <int D>
__global__ void k_exact_pass2(const float* Xp, uint32_t n, uint32_t d, uint32_t dp,
                              const uint32_t* queries, uint32_t nq, uint32_t seg_len,
                              const float* T, uint32_t* cnt, uint32_t* cand, uint32_t cap);
__global__ void k_exact_rerank(const float* Xp, uint32_t n, uint32_t d, uint32_t dp,
                               const uint32_t* queries, uint32_t nq, uint32_t k,
                               const uint32_t* cnt, const uint32_t* cand, uint32_t cap,
                               uint32_t* idx_out, float* dist_out, uint32_t* overflow,
                               uint32_t* n_overflow);
__global__ void k_exact_fallback(const float* Xp, uint32_t n, uint32_t d, uint32_t dp,
                                 const uint32_t* queries, uint32_t k, const float* T,
                                 const uint32_t* overflow, uint32_t n_overflow, double* sd,
                                 uint32_t* si, uint32_t* idx_out, float* dist_out);

}  // namespace psq
