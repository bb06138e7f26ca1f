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
__global__ void k_ctr_reduce(const unsigned long long* ctr, unsigned long long* out);
__global__ void k_join_rec_reduce(const uint32_t* rec, uint32_t n, unsigned long long* counters);
__global__ void k_sum_u32_to(const uint32_t* v, uint32_t n, unsigned long long* counters, uint32_t word);
template <int KV>
__global__ void k_sample_select(uint64_t* G, uint32_t n, uint32_t k, uint32_t m, uint64_t seed,
                                uint64_t iter, const uint32_t* roff, const uint64_t* rev,
                                uint32_t* newl, uint32_t* oldl, uint32_t* ncnt, uint64_t* slots,
                                uint32_t S, float* thr, uint32_t thr_rank);
template <int D>
__global__ void k_join(const float* Xp, uint32_t n, uint32_t d, uint32_t dp, uint32_t RS,
                       const uint32_t* newl, const uint32_t* oldl, const uint32_t* ncnt, uint32_t m,
                       uint32_t R, JoinCells jc, const float* thr, uint64_t* slots, uint8_t* touched,
                       uint32_t S, uint64_t seed, uint64_t iter, uint32_t* rec);
__global__ void k_join_defer_flags(const uint32_t* ncnt, uint32_t n, uint32_t lo, uint32_t hi,
                                   uint32_t* flag);
__global__ void k_join_defer_list(const uint32_t* flag, const uint32_t* pos, uint32_t n, uint32_t* list);
template <int P>
__global__ void k_merge_wide(uint64_t* G, uint32_t n, uint32_t k, uint64_t* slots, uint32_t S,
                             uint8_t* touched, int force, uint32_t* chg);
template <int NR, int NT>
__global__ void k_merge(uint64_t* G, uint32_t n, uint32_t k, uint64_t* slots, uint32_t S,
                        uint8_t* touched, int force, uint32_t* chg);
__global__ void k_row_counts(const uint64_t* G, uint32_t n, uint32_t k, uint32_t ko, const uint32_t* inv,
                             uint32_t* cnt);
__global__ void k_extract(const uint64_t* G, uint32_t n, uint32_t k, uint32_t ko, const uint32_t* perm,
                          const uint32_t* inv, const uint32_t* off, uint32_t* indices, float* dists);
__global__ void k_perm_inverse(const uint32_t* perm, uint32_t n, uint32_t* inv);
__global__ void k_perm_rows(const float* X, uint32_t dp, const uint32_t* perm, uint32_t n, float* Y);
__global__ void k_renumber(uint32_t* ids, size_t total, const uint32_t* inv);

// RP forest (forest_kernels.hip)
__global__ void k_forest_init(uint32_t* ids, uint32_t n, uint32_t T);
__global__ void k_forest_planes(const float* Xp, uint32_t d, uint32_t dp, uint32_t n, const uint32_t* ids,
                                const uint32_t* nstart, const uint32_t* ncount, const uint32_t* ntree,
                                uint32_t A, const uint64_t* tseeds, uint32_t level, float* normals,
                                float* offsets, float* nabs);
__global__ void k_rows_to_half(const float* X, uint32_t d, uint32_t dp, uint32_t n, _Float16* Xh,
                               uint32_t dh);
__global__ void k_forest_emit(const uint32_t* nstart, const uint32_t* ncount, const uint32_t* ntree,
                              uint32_t A, uint32_t n, const uint32_t* nl_in, const uint32_t* mode_in,
                              const uint32_t* act_off, const uint32_t* small_off, uint32_t leaf_size,
                              uint32_t cap, uint32_t level, uint32_t* x_start, uint32_t* x_count,
                              uint32_t* x_tree, uint32_t* child, uint32_t* lcnt_pos,
                              uint32_t small_base, uint32_t* sm_start, uint32_t* sm_count,
                              uint32_t* sm_tree, uint32_t* sm_level);
__global__ void k_forest_finish(const float* Xs, uint32_t ds, uint32_t dps, uint32_t RSs, uint32_t n,
                                uint32_t* ids, const uint32_t* sm_start, const uint32_t* sm_count,
                                const uint32_t* sm_tree, const uint32_t* sm_level,
                                const uint64_t* tseeds, uint32_t leaf_size, uint32_t C, uint32_t NC,
                                uint32_t* lcnt_pos, uint32_t nsmall);
constexpr uint32_t FT_TILE = 1024;  // positions per tile of the level passes
__global__ void k_ftile_count(const uint32_t* ncount, uint32_t A, uint32_t* ntile);
__global__ void k_ftile_map(const uint32_t* toff, uint32_t A, uint32_t NT, uint32_t* tnode);
__global__ void k_forest_sides_tile(const float* Xp, uint32_t d, uint32_t dp, uint32_t n,
                                    const _Float16* Xh, uint32_t dh, const uint32_t* ids,
                                    const uint32_t* nstart, const uint32_t* ncount,
                                    const uint32_t* ntree, const uint32_t* toff, const uint32_t* tnode,
                                    const uint64_t* tseeds, uint32_t level, const float* normals,
                                    const float* offsets, const float* nabs, uint64_t* flmask,
                                    uint64_t* frmask, uint32_t* tcl, uint32_t* tcr);
__global__ void k_forest_nodes_tile(const uint32_t* ncount, uint32_t A, const uint32_t* toff,
                                    const uint32_t* scl, const uint32_t* scr, uint32_t leaf_size,
                                    uint32_t cap, uint32_t level, uint32_t* nl_out, uint32_t* mode_out,
                                    uint32_t* nact, uint32_t* nsmall);
__global__ void k_forest_scatter_tile(uint32_t n, const uint32_t* ids, const uint32_t* nstart,
                                      const uint32_t* ncount, const uint32_t* ntree,
                                      const uint32_t* toff, const uint32_t* tnode,
                                      const uint32_t* nl_in, const uint32_t* mode_in,
                                      const uint64_t* flmask, const uint64_t* frmask,
                                      const uint32_t* scl, const uint32_t* scr, uint32_t* tmp);
__global__ void k_forest_copyback_tile(uint32_t n, const uint32_t* nstart, const uint32_t* ncount,
                                       const uint32_t* ntree, const uint32_t* toff,
                                       const uint32_t* tnode, const uint32_t* tmp, uint32_t* ids);
__global__ void k_leaf_flags(const uint32_t* lcnt_pos, size_t total, uint32_t* flag);
__global__ void k_leaf_compact(const uint32_t* lcnt_pos, size_t total, uint32_t n, const uint32_t* off,
                               uint32_t* g_start, uint32_t* g_count);
__global__ void k_tree_offsets(const uint32_t* off, uint32_t n, uint32_t tb, uint32_t* out);
template <int D, int KV>
__global__ void k_leaf_merge_wide(const float* Xp, uint32_t d, uint32_t dp, uint32_t RS, uint32_t n,
                                  const uint32_t* ids_t, const uint32_t* g_start, const uint32_t* g_count,
                                  uint32_t nleaves, uint64_t* G, uint32_t k, uint32_t LG);
template <int D, int KMAX, int LG, int KE>
__global__ void k_leaf_merge(const float* Xp, uint32_t d, uint32_t dp, uint32_t RS, uint32_t n,
                             const uint32_t* ids_t, const uint32_t* g_start, const uint32_t* g_count,
                             uint32_t nleaves, uint64_t* G, uint32_t k);

// precision form (bf16_kernels.hip)
template <int D>
__global__ void k_join_bf16(const float* Xp, uint32_t n, uint32_t d, uint32_t dp, uint32_t RS,
                            const uint32_t* newl, const uint32_t* oldl, const uint32_t* ncnt,
                            uint32_t m, uint32_t R, JoinCells jc, const float* thr, float slack,
                            uint64_t* slots, uint8_t* touched, uint32_t S, uint64_t seed,
                            uint64_t iter, uint32_t* rec);

// learned encoder (mlp_kernels.hip)
constexpr int MLP_N_PARAMS = 128 * 50 + 128 + 32 * 128 + 32;
__global__ void k_mlp_round(const float* w, float* wr, uint32_t count);
// pp.neighbors' cosine metric on the approximate path: unit rows as numpy
// scales them, and 1 - cos distances from the unit-row Euclidean ones.
constexpr uint32_t ROWS_D_MAX = 256;
__global__ void k_unit_rows(const float* x, uint64_t n, uint32_t d, float* out);
__global__ void k_half_square(float* dist, uint64_t count);

// Encode on the FP32 matrix cores: operands packed per lane (ENC_PACKED floats),
// ENC_WAVES wave-tiles of 16 rows per block.
constexpr int ENC_PACKED = 10752;
constexpr int ENC_WAVES = 4;
__global__ void k_encode_pack(const float* pr, float* packed);
__global__ void k_encode_mfma(const float* X, uint32_t rows, const float* packed, const float* pr,
                              float mu, float inv_scale, float* z_out);
constexpr int MLP_STEP_T = 32;     // triplets per block of k_mlp_step
constexpr int MLP_STEP_WAVES = 16;  // its waves
size_t mlp_step_smem();          // its dynamic LDS
__global__ void k_mlp_step(const float* X, const uint32_t* rows, uint32_t B, const uint32_t* pool,
                           uint32_t A, const uint32_t* knn, uint32_t kk, uint32_t ns, uint64_t seed,
                           uint64_t step, const float* pr, float mu, float inv_scale, float margin,
                           float* partial, float* loss_terms);
__global__ void k_mlp_reduce_adam(const float* partial, uint32_t nblocks, float* w, float* wr,
                                  float* m1, float* m2, float lr, float b1, float b2, float eps,
                                  float c1, float c2, float* grad_out);
__global__ void k_gather_rows(const float* src, uint32_t d, const uint32_t* idx, uint32_t rows, float* dst);
__global__ void k_moments(const float* X, uint32_t n, uint32_t d, uint32_t chunk, double* partial);

// exact kNN (exact_kernels.hip)
__global__ void k_exact_norms(const float* Xp, uint32_t n, uint32_t d, uint32_t dp, float* xn,
                              uint32_t* pmax);
size_t exm_smem(uint32_t d);  // dynamic LDS of k_exact_pass1 / k_exact_pass2
constexpr uint32_t EXM_P_TILE = 64;  // points per LDS tile of the two passes
// Query sets of 16 per wave in a filter pass: two at the fixed widths with
// lists of at most 32 (K = 0 names pass 2), else one.
__host__ __device__ constexpr uint32_t exm_qsets(int D, int K) { return D > 0 && K <= 32 ? 2u : 1u; }
// Queries per block of a filter pass (four waves).
__host__ __device__ constexpr uint32_t exm_queries(int D, int K) { return 64u * exm_qsets(D, K); }
template <int D, int K>
__global__ void k_exact_pass1(const float* Xp, const float* xn, uint32_t n, uint32_t d, uint32_t dp,
                              const uint32_t* queries, uint32_t nq, uint32_t seg_len,
                              uint32_t nseg, uint32_t k, float* partial);
__global__ void k_exact_tau(const float* partial, uint32_t nq, uint32_t L, uint32_t k, uint32_t d,
                            const uint32_t* queries, const float* xn, const uint32_t* pmax, float* T);
template <int D>
__global__ void k_exact_pass2(const float* Xp, const float* xn, uint32_t n, uint32_t d, uint32_t dp,
                              const uint32_t* queries, uint32_t nq, uint32_t seg_len,
                              const float* T, uint32_t* cnt, uint32_t* cand, uint32_t cap);
__global__ void k_exact_rerank(const float* Xp, uint32_t n, uint32_t d, uint32_t dp,
                               const uint32_t* queries, uint32_t nq, uint32_t k,
                               const uint32_t* cnt, const uint32_t* cand, uint32_t cap,
                               uint32_t* idx_out, float* dist_out, uint32_t* overflow,
                               uint32_t* n_overflow);
constexpr uint32_t EX_KBIG_MAX = 2048;  // largest k of the exact path
__global__ void k_exact_tau_big(const float* partial, uint32_t nq, uint32_t L, uint32_t k, uint32_t d,
                                const uint32_t* queries, const float* xn, const uint32_t* pmax, float* T);
__global__ void k_exact_rerank_big(const float* Xp, uint32_t n, uint32_t d, uint32_t dp,
                                   const uint32_t* queries, uint32_t nq, uint32_t k,
                                   const uint32_t* cnt, const uint32_t* cand, uint32_t cap,
                                   uint32_t pmax, uint32_t* idx_out, float* dist_out,
                                   uint32_t* overflow, uint32_t* n_overflow);
__global__ void k_iota_u32(uint32_t* q, uint32_t q0, uint32_t count);
__global__ void k_exact_fallback(const float* Xp, uint32_t n, uint32_t d, uint32_t dp,
                                 const uint32_t* queries, uint32_t k, const float* T,
                                 const uint32_t* overflow, uint32_t n_overflow, double* sd,
                                 uint32_t* si, uint32_t* idx_out, float* dist_out, const uint64_t* off,
                                 uint32_t S, double thr_scale);
template <int D, int K>
__global__ void k_bexact_pass1(const float* Xp, const float* xn, uint32_t d, uint32_t dp, const uint64_t* off,
                               const uint32_t* units, uint32_t U, uint32_t seg_len, const uint32_t* nseg_of,
                               const uint64_t* part_off, uint32_t k, uint32_t* counter, float* partial);
__global__ void k_bexact_tau(const float* partial, const uint64_t* off, uint32_t S,
                             const uint32_t* nseg_of, const uint64_t* part_off, uint64_t nq, uint32_t K,
                             uint32_t k, uint32_t d, const float* xn, const uint32_t* pmax, float* T);
template <int D>
__global__ void k_bexact_pass2(const float* Xp, const float* xn, uint32_t d, uint32_t dp, const uint64_t* off,
                               const uint32_t* units, uint32_t U, uint32_t seg_len, const float* T,
                               uint32_t* counter, uint32_t* cnt, uint32_t* cand, uint32_t cap);
// The fallback's FP32 VALU distance is within (d + 4) 2^-24 of the exact one,
// relative: the matrix-core threshold scaled by this bounds it for every true
// neighbour.
inline double exact_fallback_scale(uint64_t d) { return 1.0 + 4.0 * ((double)d + 8.0) * 5.9604644775390625e-08; }

// Level-2 partition and reverse lists (partition.hip; layout in psq_partition.hpp)
__global__ void k_part_l2_count(const uint32_t* sstart, const uint32_t* pj1, uint32_t sshift,
                                uint32_t blog, uint32_t* btot);
__global__ void k_part_l2_scatter(const uint32_t* sstart, const uint64_t* pay1, const uint32_t* pj1,
                                  uint32_t sshift, uint32_t blog, const uint32_t* bstart,
                                  uint32_t* bfill, uint64_t* pay2, uint16_t* pl2);
__global__ void k_rv_l1_count(const uint64_t* G, uint64_t nk, uint32_t sshift, uint32_t ns,
                              uint32_t* stot);
__global__ void k_rv_l1_scatter(const uint64_t* G, uint64_t nk, uint32_t k, uint64_t seed,
                                uint64_t iter, uint32_t sshift, uint32_t ns, const uint32_t* sstart,
                                uint32_t* sfill, uint64_t* pay1, uint32_t* pj1);
__global__ void k_rv_bucket_sort(const uint32_t* bstart, const uint64_t* pay2, const uint16_t* pl2,
                                 uint32_t n, uint32_t blog, uint32_t nb, uint32_t* roff,
                                 uint64_t* rev);

// Fuzzy connectivities (fuzzy_kernels.hip)
constexpr int FZ_KR = 16;          // row length held in registers by k_fuzzy_strengths
constexpr int FZ_BLOCK = 256;      // threads of the block-per-row kernels
constexpr int FZ_MID_CAP = 8192;   // longest combined row sorted in LDS
constexpr int FZ_EMIT_LANES = 16;  // lanes per row in k_fuzzy_emit
constexpr int FZ_BUCKET_LOG = 8;       // 256 rows per bucket
constexpr int FZ_BUCKET_ROWS = 1 << FZ_BUCKET_LOG;
constexpr int FZ_BCAP = 9216;          // bucket entries counting-sorted in LDS
constexpr int FZ_MB = 1024;            // threads of k_fz_bucket_merge
__global__ void k_fuzzy_check(const uint32_t* indptr, uint32_t n, const uint32_t* indices, uint64_t m,
                              uint32_t* bad);
__global__ void k_fuzzy_strengths(const uint32_t* indptr, const float* dists, uint32_t n,
                                  uint64_t* pay0);
template <int SBM>
__global__ void k_fz_l1_count(const uint32_t* indices, uint64_t m, uint32_t sshift, uint32_t ns,
                              uint32_t* stot);
template <int SBM>
__global__ void k_fz_l1_scatter(const uint64_t* pay0, const uint32_t* indices, uint64_t m,
                                uint32_t sshift, uint32_t ns, const uint32_t* sstart,
                                uint32_t* sfill, uint64_t* pay1, uint32_t* pj1);
__global__ void k_fz_bucket_merge(const uint32_t* indptr, const uint32_t* indices,
                                  const uint64_t* pay0, uint32_t n, uint32_t nb,
                                  const uint32_t* bstart, const uint64_t* pay2, const uint16_t* pl2,
                                  uint32_t* t_ptr, uint64_t* comb, uint32_t* olen, uint32_t* big,
                                  uint32_t* nbig);
__global__ void k_fuzzy_merge_mid(const uint32_t* indptr, const uint32_t* t_ptr, uint64_t* comb,
                                  uint32_t* olen, const uint32_t* big, const uint32_t* nbig,
                                  uint32_t* giant, uint32_t* ngiant);
__global__ void k_fuzzy_row_len(const uint32_t* indptr, const uint32_t* t_ptr, const uint32_t* rows,
                                uint32_t nrows, uint32_t* len);
__global__ void k_fuzzy_gather_pad(const uint32_t* indptr, const uint32_t* t_ptr, const uint32_t* rows,
                                   uint32_t nrows, uint32_t P, const uint64_t* comb, uint64_t* tmp);
__global__ void k_fuzzy_bitonic_step(uint64_t* tmp, uint32_t nrows, uint32_t P, uint32_t size,
                                     uint32_t stride);
__global__ void k_fuzzy_fold_rows(const uint32_t* indptr, const uint32_t* t_ptr, const uint32_t* rows,
                                  uint32_t nrows, uint32_t P, const uint64_t* tmp, uint64_t* comb,
                                  uint32_t* olen);
__global__ void k_fuzzy_total(const uint32_t* olen, uint32_t n, unsigned long long* total);
__global__ void k_fuzzy_emit(const uint32_t* indptr, const uint32_t* t_ptr, uint32_t n,
                             const uint64_t* comb, const uint32_t* optr, uint32_t* out_ix,
                             float* out_w);

// Exact neighbour graph of neighbors.rs:101-241 (graph_exact.hip)
constexpr uint32_t NBR_LDS_K = 128;    // heaps of at most this many entries live in LDS
constexpr uint32_t NBR_K_MAX = 16384;  // largest min(n_neighbors, n_obs): one LDS sort per row
__global__ void k_nbr_row_norms(const float* X, uint32_t n, uint32_t d, float* rs);
template <int D, int M>
__global__ void k_nbr_heap(const float* X, uint32_t n, uint32_t d, uint32_t k, const float* rs,
                           float* hd, uint32_t* hi, uint32_t lds_heap);
__global__ void k_nbr_sort64(const float* hd, const uint32_t* hi, uint32_t n, uint32_t k, int metric,
                             uint32_t* out_idx, float* out_d);
__global__ void k_nbr_sort_lds(const float* hd, const uint32_t* hi, uint32_t n, uint32_t k, uint32_t P,
                               int metric, uint32_t* out_idx, float* out_d);
__global__ void k_nbr_indptr(uint32_t* indptr, uint32_t n, uint32_t per);
__global__ void k_single_point(const float* X, uint32_t d, uint32_t* indptr, uint32_t* indices,
                               float* dists);

}  // namespace psq
