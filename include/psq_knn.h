/* psq_knn.h — C-ABI of the MI355X neighbour-graph library (libpsq_knn.so).
 *
 * Drop-in boundary for Polariseq's neighbour-graph step. Each entry point names
 * the reference interface it replaces (paths relative to the reference root).
 *
 * Rules, for every function:
 *  - Return 0 (PSQ_OK) on success, a negative PSQ_E* code on failure. Nothing
 *    aborts or throws across this boundary. psq_last_error() gives the message
 *    of the last failure on the calling thread.
 *  - Host-pointer entry points (no "_device" suffix) take caller-owned host
 *    buffers and copy to and from the GPU. Entries past nnz are not touched.
 *  - "_device" entry points take device pointers on the current HIP device and
 *    run on `stream` (NULL = the default stream). They return after the
 *    results are complete in device memory.
 *  - The library owns its device working set. It grows as needed, is kept
 *    between calls, and is released by psq_release_workspace(). Calls are
 *    reentrant: every entry point that uses a device's working set holds that
 *    device's lock for the whole call, so calls from several threads on one
 *    device run one after the other (the reference's calls are reentrant too).
 *    The exception is psq_knn_exact_device_noalloc (see there).
 *  - Graph layout (crates/polariseq-core/src/backend.rs:37-46): CSR with indptr
 *    u32[n_obs+1], indices u32[nnz], dists f32[nnz]. Rows are ascending by
 *    distance, ties by index. Distances are Euclidean (sqrt of the squared
 *    distance). No self-loops, no duplicates, at most k per row.
 */
#ifndef PSQ_KNN_H
#define PSQ_KNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSQ_OK 0
#define PSQ_E_INVALID (-1)     /* bad argument: empty data, k == 0, short outputs, ... */
#define PSQ_E_UNSUPPORTED (-2) /* valid, but outside what this build supports */
#define PSQ_E_DEVICE (-3)      /* a HIP call failed (message names it) */
#define PSQ_E_NOMEM (-4)       /* device memory (or the debug budget) exhausted */

#define PSQ_PRECISION_FP32 0        /* plain form: exact f32 sq_dist of the reference */
#define PSQ_PRECISION_BF16_RERANK 1 /* BF16 MFMA tiles screen the join's pairs, screened pairs
                                       get the exact f32 distance; rows keep their top 2k */

#define PSQ_INIT_RP_FOREST 0 /* random-projection forest (the reference's start) */
#define PSQ_INIT_RANDOM 1    /* random fill only */
#define PSQ_INIT_LEARNED 2   /* candidates from a learned encoder's codes (model != NULL) */

/* Mirrors NndescentOptions, crates/polariseq-core/src/nndescent.rs:41-76.
 * 0 means "auto" for the sized fields, resolved as in nndescent.rs:722-752. */
typedef struct psq_nndescent_opts {
  uint64_t k;              /* neighbours per node (>= 1) */
  uint64_t max_candidates; /* 0 -> min(60, 2k) */
  uint64_t max_iterations; /* 0 -> max(5, round(log2 n)) */
  double delta;            /* stop when changed <= floor(delta * n * heap_k); default 0.001 */
  uint64_t n_trees;        /* 0 -> min(48, 8 + round(sqrt(n)/10)) */
  uint64_t leaf_size;      /* 0 -> max(10, k) */
  uint64_t gather_max;     /* accepted for compatibility; a memory choice that never changes the graph */
  uint64_t seed;
  int32_t precision; /* PSQ_PRECISION_* */
  int32_t init;      /* PSQ_INIT_* */
  const void* model; /* learned encoder (PSQ_INIT_LEARNED), see psq_encoder_* */
  int32_t device;    /* HIP device ordinal for host-pointer entries */
  uint32_t slots;    /* proposal slots per row per iteration (0 -> auto, power of two >= 4 heap_k) */
  uint32_t heap_k;   /* entries each row keeps while iterating, >= k (0 -> k for FP32,
                        min(2k, 64) for BF16_RERANK); the output holds each row's first k */
  uint32_t reorder;  /* renumber the points in the first tree's leaf order while iterating
                        (0 -> on for BF16_RERANK with a forest start, 1 on, 2 off); the graph
                        is returned in the caller's numbering either way */
} psq_nndescent_opts;

/* Defaults of NndescentOptions::default (nndescent.rs:63-76) with k = 15. */
void psq_nndescent_opts_default(psq_nndescent_opts* o);

/* crates/polariseq-core/src/nndescent.rs:648-650: max(1, min(k, n_obs - 1)). */
uint64_t psq_effective_k(uint64_t n_obs, uint64_t k);

/* Replaces nndescent_into (crates/polariseq-core/src/nndescent.rs:699-707) and,
 * through the Python binding, nndescent_dense (crates/polariseq-python/src/
 * lib.rs:507-566). data is f32[n_obs][n_dim] on the host. indptr_len must be
 * n_obs + 1; indices and dists must hold cap >= n_obs * psq_effective_k(n_obs, k)
 * entries. Writes *nnz_out. Errors instead of the reference's panics
 * (nndescent.rs:708-719). */
int psq_nndescent_into(const float* data, uint64_t n_obs, uint64_t n_dim,
                       const psq_nndescent_opts* opts, uint32_t* indptr, uint64_t indptr_len,
                       uint32_t* indices, float* dists, uint64_t cap, uint64_t* nnz_out);

/* The same on device pointers (embedding resident in HBM, graph left in HBM). */
int psq_nndescent_device(const float* d_data, uint64_t n_obs, uint64_t n_dim,
                         const psq_nndescent_opts* opts, uint32_t* d_indptr, uint64_t indptr_len,
                         uint32_t* d_indices, float* d_dists, uint64_t cap, uint64_t* nnz_out,
                         void* stream);

/* Exact kNN, self excluded, ordered by (FP64 distance, index): the order of
 * by_distance / knn_exact (crates/polariseq-core/src/doublets.rs:544-651).
 * queries: u32 indices into data, or NULL for all n points (n_q must then be n).
 * Outputs idx_out u32[n_q][k] and dist_out f32[n_q][k] (Euclidean, from the FP64
 * squared distance). Rows with fewer than k other points are padded with index
 * 0xFFFFFFFF and distance +inf. k <= 2048 (k > 64 covers Scrublet's
 * k_adj = round(0.5 sqrt(n_cells)) * 3, doublets.rs:1083-1090); larger k returns
 * PSQ_E_UNSUPPORTED. */
int psq_knn_exact(const float* data, uint64_t n, uint64_t d, const uint32_t* queries, uint64_t n_q,
                  uint64_t k, uint32_t* idx_out, float* dist_out, int32_t device);
int psq_knn_exact_device(const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries,
                         uint64_t n_q, uint64_t k, uint32_t* d_idx_out, float* d_dist_out,
                         void* stream);

/* No counterpart in the reference: psq_knn_exact_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It calls no hipMalloc, hipFree or allocating
 * library routine: its buffers live in a static device array of
 * PSQ_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until the
 * process ends. Results equal psq_knn_exact_device's bit for bit. k <= 64
 * (larger k returns PSQ_E_UNSUPPORTED); inputs whose buffers do not fit the
 * array return PSQ_E_INVALID before any work. They need at least
 * n * round_up(d, 32) * 4 + n_q * (524 + 4 K + 12 k) bytes (K the smallest of
 * 16, 32, 64 that is >= k; each term rounded up to 256 bytes): 64 MiB holds,
 * e.g., 100,000 points of 50 dims with 40,000 queries at k = 15. Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU race, so the
 * caller serialises them (one stream, or its own lock). */
#define PSQ_NOALLOC_SCRATCH_BYTES (64ull << 20)
int psq_knn_exact_device_noalloc(const float* d_data, uint64_t n, uint64_t d,
                                 const uint32_t* d_queries, uint64_t n_q, uint64_t k,
                                 uint32_t* d_idx_out, float* d_dist_out, void* stream);

/* Batched exact kNN: n_samples independent samples stored as contiguous row
 * ranges [offsets[s], offsets[s+1]) of data (offsets is a HOST array of
 * n_samples + 1 entries, offsets[0] = 0, offsets[n_samples] = n_total). Every
 * point is a query against the points of its own sample, self excluded, in the
 * order of psq_knn_exact; indices returned are global row indices. All samples
 * run in one persistent launch per pass; tile shapes come from an autotuner
 * that times candidate shapes once per shape class and caches the winner.
 * Reference analogue: Scrublet's per-sample exact kNN
 * (crates/polariseq-core/src/doublets.rs:557-651, 1085-1113, 1366-1374). */
int psq_knn_exact_batched(const float* data, uint64_t n_total, uint64_t d, const uint64_t* offsets,
                          uint64_t n_samples, uint64_t k, uint32_t* idx_out, float* dist_out,
                          int32_t device);
int psq_knn_exact_batched_device(const float* d_data, uint64_t n_total, uint64_t d,
                                 const uint64_t* offsets, uint64_t n_samples, uint64_t k,
                                 uint32_t* d_idx_out, float* d_dist_out, void* stream);
typedef struct psq_batched_report {
  uint32_t block, tile, seg_len; /* the tile shape used */
  uint32_t tuned_now;            /* 1 when this call ran the autotuner */
  uint32_t configs_tried, units, overflow_queries, pad;
  double tune_ms, ms_pass1, ms_pass2, ms_rerank, ms_total;
  double pairs; /* Σ n_s^2 distance pairs of the samples */
} psq_batched_report;
int psq_batched_last_report(psq_batched_report* out);

/* Replaces neighbors() / neighbors_dense (crates/polariseq-core/src/neighbors.rs:101-241,
 * crates/polariseq-python/src/lib.rs:474-501), with the reference's results bit for bit:
 * k_with_self = min(n_neighbors, n_obs); each query keeps a bounded max-heap of
 * k_with_self entries over every point in index order (a point replaces the root only
 * when strictly closer), the heap array is sorted stably by distance and its entry 0 is
 * dropped as self (so a duplicate of the query can take that place and the query stay in
 * its own row). Distances are the reference's f32 values: metric 0 = euclidean, the sqrt
 * of the 8-lane sq_dist; 1 = cosine, 1 - dot/(sqrt(dot(a,a)) sqrt(dot(b,b))), and 1.0
 * when that denominator is 0. Each row holds k_with_self - 1 entries; capacity
 * n_obs * (k_with_self - 1). k_with_self > 16384 returns PSQ_E_UNSUPPORTED. Rows of
 * distances that contain NaN are ordered with NaN last (the reference's order is
 * unspecified there). */
int psq_neighbors_exact(const float* embedding, uint64_t n_obs, uint64_t n_dim,
                        uint64_t n_neighbors, int32_t metric, uint32_t* indptr, uint32_t* indices,
                        float* dists, uint64_t cap, uint64_t* nnz_out, int32_t device);
/* The same on device pointers: embedding f32[n_obs][n_dim], indptr u32[n_obs + 1],
 * indices / dists of cap entries, on the current device and `stream`. */
int psq_neighbors_exact_device(const float* d_embedding, uint64_t n_obs, uint64_t n_dim,
                               uint64_t n_neighbors, int32_t metric, uint32_t* d_indptr,
                               uint32_t* d_indices, float* d_dists, uint64_t cap,
                               uint64_t* nnz_out, void* stream);

/* pp.neighbors' cosine metric on the approximate path
 * (python/polariseq/__init__.py:2540-2552): NN-Descent searches unit rows, on which
 * Euclidean order is cosine order, and the distances become 1 - cos.
 * psq_unit_rows_device: d_out[i] = d_x[i] / max(|d_x[i]|, 1e-12), bit for bit
 * as numpy's float32 `emb / np.maximum(np.linalg.norm(emb, axis=1, keepdims=True),
 * 1e-12)` (numpy's pairwise summation order for the norm); n_dim <= 256.
 * psq_cosine_distances_device: d_dist[e] = float32(float64(d_dist[e])^2 / 2), as
 * numpy's `(d.astype(np.float64) ** 2 / 2.0).astype(np.float32)`. In place. */
int psq_unit_rows_device(const float* d_x, uint64_t n, uint64_t d, float* d_out, void* stream);
int psq_cosine_distances_device(float* d_dist, uint64_t count, void* stream);

/* PCA of a CSR matrix (n_obs x n_features; indptr u64[n_obs + 1], indices u32,
 * data f32). Replaces pca_csr_into / pca_scaled_blocks_into
 * (crates/polariseq-core/src/pca.rs:255-290, 396-431) as bound by pca_sparse and
 * pca_sparse_scaled (crates/polariseq-python/src/lib.rs:395-467).
 * scale = 0: centre when center != 0 (implicitly; the matrix is never changed).
 * scale = 1: the pp.scale view (ddof-1 sd, 0 -> 1, values clipped to
 *   +-max_value; max_value = +inf for none, else >= 0), always centred.
 * method: PSQ_PCA_GRAM, the exact Gram route (pca.rs:487-535), n_features <=
 *   32768; PSQ_PCA_RANDOMIZED, randomized block Krylov SVD with implicit
 *   centring (rsvd.rs:476-566, pca.rs:273-286), configured by oversample,
 *   n_iter (the Krylov depth) and seed (the reference's ChaCha8 Gaussian
 *   sketch, rsvd.rs:414-435), its basis (n_components + oversample)(n_iter + 1)
 *   columns (capped at min(n_obs, n_features)) at most 1024, n_obs < 2^31;
 *   PSQ_PCA_AUTO, the Gram route up to 32768 genes and the randomized route
 *   above. The reference takes the randomized route above 4096 genes
 *   (pca.rs:201); this build stays exact up to 32768 (the result meets the
 *   reference's acceptance bounds against the exact SVD, and more; the Gram
 *   product costs n_obs n_features^2 flops, so wide matrices of many cells are
 *   slow there: PSQ_PCA_RANDOMIZED asks for the reference's route instead).
 * Outputs: embedding f32[n_obs][n_components] (columns signed so that each
 * column's first largest-magnitude entry is positive), variance_ratio
 * f32[n_components] (Gram: max(eigenvalue, 0) / trace of the centred Gram;
 * randomized: s_j^2 / ||A - 1 mu^T||_F^2), and, if components is not NULL,
 * the principal axes f32[n_features][n_components], with the embedding's
 * column signs (Gram: the centred matrix times the axes is the embedding;
 * randomized: the sketch's right singular vectors B^T W / s).
 * n_obs or n_features 0, or n_components 0 or above min(n_obs, n_features),
 * PSQ_E_INVALID with the reference's message. A run repeats bit for bit.
 * The _device forms take device pointers and run on the given stream of the
 * current device. psq_pca_csr and psq_pca_csr_device are the _ex forms with
 * PSQ_PCA_AUTO and the reference's default options (pca.rs:73-83: oversample
 * 10, n_iter 1, seed 0). Unlike the neighbour-graph entries, PCA allocates its
 * device buffers per call (freed before it returns, outside the
 * psq_set_device_budget cap); the rocBLAS handle it creates on a device's
 * first call lives until the process ends. */
#define PSQ_PCA_AUTO 0
#define PSQ_PCA_GRAM 1
#define PSQ_PCA_RANDOMIZED 2
int psq_pca_csr(const uint64_t* indptr, const uint32_t* indices, const float* data, uint64_t n_obs,
                uint64_t n_features, uint64_t n_components, int32_t center, int32_t scale,
                double max_value, float* embedding, float* variance_ratio, float* components,
                int32_t device);
int psq_pca_csr_device(const uint64_t* d_indptr, const uint32_t* d_indices, const float* d_data,
                       uint64_t n_obs, uint64_t n_features, uint64_t n_components, int32_t center,
                       int32_t scale, double max_value, float* d_embedding, float* d_variance_ratio,
                       float* d_components, void* stream);
int psq_pca_csr_ex(const uint64_t* indptr, const uint32_t* indices, const float* data, uint64_t n_obs,
                   uint64_t n_features, uint64_t n_components, int32_t center, int32_t scale,
                   double max_value, int32_t method, uint64_t oversample, uint64_t n_iter,
                   uint64_t seed, float* embedding, float* variance_ratio, float* components,
                   int32_t device);
int psq_pca_csr_device_ex(const uint64_t* d_indptr, const uint32_t* d_indices, const float* d_data,
                          uint64_t n_obs, uint64_t n_features, uint64_t n_components, int32_t center,
                          int32_t scale, double max_value, int32_t method, uint64_t oversample,
                          uint64_t n_iter, uint64_t seed, float* d_embedding, float* d_variance_ratio,
                          float* d_components, void* stream);
/* Phases of this thread's last successful PCA call, from HIP events on its
 * stream. Gram route: the scaled view, the Gram (densified chunks, column sums
 * and the DGEMM products; ms_gemm is those launches alone), the eigensolver,
 * and projection + signs + ratios. Randomized route: ms_gram holds the column
 * sums and every pass building the Krylov basis and B^T (ms_gemm their DGEMM
 * launches), ms_eigen the small eigenproblem, ms_project U, the outputs and
 * signs; iterations counts the passes over the matrix and block the basis
 * width. */
typedef struct psq_pca_report {
  uint64_t iterations; /* Gram: block products with C in the eigensolver; randomized: passes */
  uint64_t block;      /* Gram: subspace block b; randomized: Krylov basis width */
  uint64_t chunks;     /* dense row chunks of a pass */
  uint64_t rows_per_chunk;
  double ms_scale, ms_gram, ms_gemm, ms_eigen, ms_project, ms_total;
  uint64_t method; /* PSQ_PCA_GRAM or PSQ_PCA_RANDOMIZED */
} psq_pca_report;
int psq_pca_last_report(psq_pca_report* out);

/* ── learned encoder (no reference counterpart; used by PSQ_INIT_LEARNED) ──
 * A 50 -> 128 -> 32 MLP (GELU, BF16-rounded weights, FP32 accumulation),
 * trained with a triplet margin loss on exact-kNN triplets of a subset. */
typedef struct psq_encoder psq_encoder;
typedef struct psq_encoder_train_opts {
  uint64_t subset; /* cells in the training subset (0 -> 200000), evenly spaced */
  uint64_t k_pos;  /* positives drawn from the exact top-k_pos within the subset (0 -> 15) */
  uint64_t batch;  /* triplets per step (0 -> 8192) */
  uint64_t steps;  /* Adam steps (0 -> 100) */
  float lr;        /* 0 -> 3e-3 */
  float margin;    /* triplet margin on squared code distances (0 -> 0.01) */
  uint64_t seed;
  uint64_t anchors; /* anchor pool: evenly spaced subset cells whose exact top-k_pos is
                       mined (0 -> 4096); anchors are drawn from it, negatives from the subset */
} psq_encoder_train_opts;
int psq_encoder_create(psq_encoder** out, uint64_t seed, int32_t device);
void psq_encoder_destroy(psq_encoder* e);
/* Back to the untrained state of psq_encoder_create(seed): initial weights, zero Adam
 * moments, no standardisation; the device buffers are kept. */
int psq_encoder_reset(psq_encoder* e, uint64_t seed);
int psq_encoder_train(psq_encoder* e, const float* d_data, uint64_t n, uint64_t d,
                      const psq_encoder_train_opts* o, void* stream);
int psq_encoder_encode(const psq_encoder* e, const float* d_data, uint64_t n, uint64_t d,
                       float* d_codes, void* stream);
int psq_encoder_params(const psq_encoder* e, float* host_params, uint64_t count, float* mu_scale);
int psq_encoder_set_params(psq_encoder* e, const float* host_params, uint64_t count, float mu,
                           float inv_scale);
int psq_encoder_grad(const psq_encoder* e, const float* d_data, uint64_t n, uint64_t d,
                     const uint32_t* rows, uint64_t B, float margin, float* host_grad,
                     double* loss_out);
int psq_encoder_info(const psq_encoder* e, double* last_loss, double* train_ms, uint64_t* steps);
/* HIP-event time of the last training's mining part (subset, moments, exact kNN of the
 * anchor pool); train_ms includes it. */
int psq_encoder_mine_ms(const psq_encoder* e, double* mine_ms);

/* Fuzzy connectivities of a kNN distance graph: the symmetric CSR that clustering
 * and UMAP share (scanpy's obsp['connectivities']), replacing
 * fuzzy_connectivities_from (crates/polariseq-core/src/umap.rs:130-139). Each
 * row's membership strengths come from the bandwidth search of umap.rs:551-608;
 * row i of the result is the fuzzy union w + w' - w * w' of its own entries with
 * the entries of the rows that name i (umap.rs:488-537), sorted by column, with
 * unions <= 1e-5 dropped. The input is a CSR with indptr[0] == 0, non-decreasing
 * row pointers and every index < n_obs; the result has at most 2 * indptr[n_obs]
 * entries. A result of 2^32 entries or more fails with PSQ_E_UNSUPPORTED, as the
 * reference binding does (crates/polariseq-python/src/lib.rs:616-619). */

/* Device form: inputs and outputs on the current device. The outputs hold
 * n_obs + 1 row pointers and cap entries; cap >= 2 * indptr[n_obs] always
 * suffices. *nnz_out is set also when cap is too small (then PSQ_E_INVALID and
 * nothing is written). */
int psq_fuzzy_connectivities_device(const uint32_t* d_indptr, const uint32_t* d_indices,
                                    const float* d_dists, uint64_t n_obs, uint32_t* d_out_indptr,
                                    uint32_t* d_out_indices, float* d_out_weights, uint64_t cap,
                                    uint64_t* nnz_out, void* stream);

/* Host form, mirroring FuzzyPlan (umap.rs:327-460) as fuzzy_connectivities_into
 * uses it (crates/polariseq-python/src/lib.rs:590-642): psq_fuzzy_plan_new
 * computes the graph on the GPU and reports its nnz (FuzzyPlan::new +
 * row_lens), the caller allocates, psq_fuzzy_plan_fill copies the graph into
 * its arrays (FuzzyPlan::fill), psq_fuzzy_plan_free releases the plan's device
 * memory. nnz_in is the length of indices and dists (>= indptr[n_obs]). */
typedef struct psq_fuzzy_plan psq_fuzzy_plan;
int psq_fuzzy_plan_new(const uint32_t* indptr, uint64_t indptr_len, const uint32_t* indices,
                       const float* dists, uint64_t nnz_in, uint64_t n_obs, int32_t device,
                       psq_fuzzy_plan** plan, uint64_t* nnz_out);
int psq_fuzzy_plan_fill(const psq_fuzzy_plan* plan, uint32_t* indptr, uint64_t indptr_len,
                        uint32_t* indices, float* weights, uint64_t cap);
void psq_fuzzy_plan_free(psq_fuzzy_plan* plan);

/* Per-run statistics of the last NN-Descent call on this thread (for reports). */
typedef struct psq_nndescent_stats {
  uint64_t n_obs, n_dim, k, max_candidates, max_iterations, n_trees, leaf_size, slots, stop_at;
  uint64_t iterations;
  uint64_t heap_k;        /* entries kept per row while iterating */
  uint64_t reordered;     /* 1 when the points were renumbered while iterating */
  uint64_t changed[64];   /* per iteration */
  uint64_t pairs[64];     /* distance pairs evaluated in the join, per iteration */
  uint64_t cand_rows[64]; /* Σ_i (|new_i| + |old_i|), the join's row gathers, per iteration */
  uint64_t exact_pairs[64]; /* pairs whose exact f32 distance the join computed (all pairs in
                               the FP32 form, the BF16 screen's survivors in BF16_RERANK) */
  uint64_t proposals[64]; /* (row, neighbour) proposals the join sent to the buckets */
  double ms_init, ms_iterations, ms_extract, ms_total;
  double ms_join[64]; /* join kernel time per iteration (HIP events) */
} psq_nndescent_stats;
int psq_last_stats(psq_nndescent_stats* out);

/* Message of the last failure on this thread ("" if none). */
const char* psq_last_error(void);

/* Frees the cached device working set of the current device. */
int psq_release_workspace(void);

/* Debug cap on the library's device allocations, in bytes (0 = none): an
 * allocation past it fails with PSQ_E_NOMEM, so out-of-memory handling can be
 * tested from the host. */
void psq_set_device_budget(uint64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* PSQ_KNN_H */
