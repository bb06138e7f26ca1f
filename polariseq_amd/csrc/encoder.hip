// C-ABI of the learned encoder: create, train on exact-kNN triplets, encode.
// No reference counterpart: the reference has no learned component. The codes
// seed NN-Descent (PSQ_INIT_LEARNED) in place of the forest's start on the
// raw embedding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psq_knn.h"
#include "psq_kernels.hpp"

using namespace psq;

namespace psq_host {
int fail(int code, const char* fmt, ...);
int knn_exact_device_impl(const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries,
                          uint64_t nq, uint64_t k, uint32_t* d_idx, float* d_dist, hipStream_t s);
int lock_device(std::unique_lock<std::recursive_mutex>* out);
}  // namespace psq_host

using psq_host::fail;

#define EHIPC(call)                                                                         \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(e_ == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE, "%s failed: %s", \
                  #call, hipGetErrorString(e_));                                            \
  } while (0)
#define ETRY(expr)                  \
  do {                              \
    int rc_ = (expr);               \
    if (rc_ != PSQ_OK) return rc_;  \
  } while (0)

struct psq_encoder {
  int device = 0;
  float mu = 0.0f, inv_scale = 1.0f;
  bool trained = false;
  float* w = nullptr;   // FP32 master weights
  float* wr = nullptr;  // BF16-rounded copy used by the products
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* packed = nullptr;  // wr laid out as the encode kernel's per-lane operands
  uint64_t steps_done = 0;
  double last_loss = 0.0;
  double train_ms = 0.0;
  double mine_ms = 0.0;  // subset, moments and the exact kNN of the anchor pool
  // training buffers, grown on demand and kept until destroy
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  Buf sub, x, pool, knn, knd, partial, lt, mom, rows, grad;
};

namespace {

constexpr uint32_t NP = MLP_N_PARAMS;

// Deterministic initial weights: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from the
// splitmix hash of (seed, parameter index); biases zero.
void init_params(std::vector<float>& p, uint64_t seed) {
  p.assign(NP, 0.0f);
  auto u = [&](uint64_t i) {
    return (float)((splitmix(seed ^ splitmix(i + 0x77ull)) >> 40) / (double)(1u << 24) * 2.0 - 1.0);
  };
  const float a1 = 1.0f / std::sqrt(50.0f), a2 = 1.0f / std::sqrt(128.0f);
  for (uint32_t i = 0; i < 128 * 50; ++i) p[i] = a1 * u(i);
  for (uint32_t i = 0; i < 32 * 128; ++i) p[128 * 50 + 128 + i] = a2 * u(100000 + i);
}

unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
int ebuf(psq_encoder::Buf& b, size_t count, T** out) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (b.bytes < bytes) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    const hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess)
      return fail(e == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE,
                  "encoder buffer of %zu bytes: %s", bytes, hipGetErrorString(e));
    b.bytes = bytes;
  }
  *out = static_cast<T*>(b.p);
  return PSQ_OK;
}

void free_bufs(psq_encoder* e) {
  for (psq_encoder::Buf* b : {&e->sub, &e->x, &e->pool, &e->knn, &e->knd, &e->partial, &e->lt,
                              &e->mom, &e->rows, &e->grad}) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->bytes = 0;
  }
}

}  // namespace

extern "C" {

int psq_encoder_create(psq_encoder** out, uint64_t seed, int32_t device) {
  if (!out) return fail(PSQ_E_INVALID, "out is NULL");
  int count = 0;
  EHIPC(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(PSQ_E_INVALID, "device %d out of range", device);
  EHIPC(hipSetDevice(device));
  psq_encoder* e = new psq_encoder();
  e->device = device;
  std::vector<float> p;
  init_params(p, seed);
  hipError_t rc = hipMalloc(&e->w, NP * sizeof(float));
  if (rc == hipSuccess) rc = hipMalloc(&e->wr, NP * sizeof(float));
  if (rc == hipSuccess) rc = hipMalloc(&e->m1, NP * sizeof(float));
  if (rc == hipSuccess) rc = hipMalloc(&e->m2, NP * sizeof(float));
  if (rc == hipSuccess) rc = hipMalloc(&e->packed, ENC_PACKED * sizeof(float));
  if (rc == hipSuccess) rc = hipMemcpy(e->w, p.data(), NP * sizeof(float), hipMemcpyHostToDevice);
  if (rc == hipSuccess) rc = hipMemset(e->m1, 0, NP * sizeof(float));
  if (rc == hipSuccess) rc = hipMemset(e->m2, 0, NP * sizeof(float));
  if (rc != hipSuccess) {
    (void)hipFree(e->w);
    (void)hipFree(e->wr);
    (void)hipFree(e->m1);
    (void)hipFree(e->m2);
    (void)hipFree(e->packed);
    delete e;
    return fail(PSQ_E_DEVICE, "encoder allocation failed: %s", hipGetErrorString(rc));
  }
  hipLaunchKernelGGL(k_mlp_round, dim3(blocks_for(NP, 256)), dim3(256), 0, 0, e->w, e->wr, NP);
  EHIPC(hipGetLastError());
  EHIPC(hipDeviceSynchronize());
  *out = e;
  return PSQ_OK;
}

int psq_encoder_reset(psq_encoder* e, uint64_t seed) {
  if (!e) return fail(PSQ_E_INVALID, "NULL encoder");
  EHIPC(hipSetDevice(e->device));
  std::vector<float> p;
  init_params(p, seed);
  EHIPC(hipMemcpy(e->w, p.data(), NP * sizeof(float), hipMemcpyHostToDevice));
  EHIPC(hipMemset(e->m1, 0, NP * sizeof(float)));
  EHIPC(hipMemset(e->m2, 0, NP * sizeof(float)));
  hipLaunchKernelGGL(k_mlp_round, dim3(blocks_for(NP, 256)), dim3(256), 0, 0, e->w, e->wr, NP);
  EHIPC(hipGetLastError());
  EHIPC(hipDeviceSynchronize());
  e->mu = 0.0f;
  e->inv_scale = 1.0f;
  e->trained = false;
  e->steps_done = 0;
  e->last_loss = e->train_ms = e->mine_ms = 0.0;
  return PSQ_OK;
}

void psq_encoder_destroy(psq_encoder* e) {
  if (!e) return;
  free_bufs(e);
  (void)hipFree(e->w);
  (void)hipFree(e->wr);
  (void)hipFree(e->m1);
  (void)hipFree(e->m2);
  (void)hipFree(e->packed);
  delete e;
}

int psq_encoder_params(const psq_encoder* e, float* host_params, uint64_t count, float* mu_scale) {
  if (!e || !host_params || count != NP) return fail(PSQ_E_INVALID, "need %u parameters", NP);
  EHIPC(hipMemcpy(host_params, e->w, NP * sizeof(float), hipMemcpyDeviceToHost));
  if (mu_scale) {
    mu_scale[0] = e->mu;
    mu_scale[1] = e->inv_scale;
  }
  return PSQ_OK;
}

int psq_encoder_set_params(psq_encoder* e, const float* host_params, uint64_t count, float mu,
                           float inv_scale) {
  if (!e || !host_params || count != NP) return fail(PSQ_E_INVALID, "need %u parameters", NP);
  EHIPC(hipMemcpy(e->w, host_params, NP * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_round, dim3(blocks_for(NP, 256)), dim3(256), 0, 0, e->w, e->wr, NP);
  EHIPC(hipGetLastError());
  EHIPC(hipDeviceSynchronize());
  e->mu = mu;
  e->inv_scale = inv_scale;
  e->trained = true;
  return PSQ_OK;
}

// Loss and gradient of one batch of triplets given as host row indices
// [anchors | positives | negatives] into d_data (n x 50, device). For tests.
int psq_encoder_grad(const psq_encoder* e, const float* d_data, uint64_t n, uint64_t d,
                     const uint32_t* rows, uint64_t B, float margin, float* host_grad,
                     double* loss_out) {
  if (!e || !d_data || !rows || !host_grad || !loss_out) return fail(PSQ_E_INVALID, "NULL argument");
  if (d != 50) return fail(PSQ_E_UNSUPPORTED, "the encoder takes 50 dimensions");
  if (B == 0) return fail(PSQ_E_INVALID, "empty batch");
  for (uint64_t i = 0; i < 3 * B; ++i)
    if (rows[i] >= n) return fail(PSQ_E_INVALID, "row %llu out of range", (unsigned long long)i);
  psq_encoder* me = const_cast<psq_encoder*>(e);
  const uint32_t nb = blocks_for(B, MLP_STEP_T);
  uint32_t* d_rows;
  float *lt, *partial, *grad;
  ETRY(ebuf(me->rows, 3 * B, &d_rows));
  ETRY(ebuf(me->lt, B, &lt));
  ETRY(ebuf(me->partial, (size_t)nb * NP, &partial));
  ETRY(ebuf(me->grad, NP, &grad));
  EHIPC(hipMemcpy(d_rows, rows, 3 * B * 4, hipMemcpyHostToDevice));
  EHIPC(hipFuncSetAttribute((const void*)k_mlp_step, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)mlp_step_smem()));
  hipLaunchKernelGGL(k_mlp_step, dim3(nb), dim3(64 * MLP_STEP_WAVES), mlp_step_smem(), 0, d_data, d_rows, (uint32_t)B, nullptr, 0u,
                     nullptr, 0u, 0u, 0ull, 0ull, e->wr, e->mu, e->inv_scale, margin, partial, lt);
  EHIPC(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(blocks_for(NP, 32)), dim3(1024), 0, 0, partial, nb,
                     nullptr, nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, grad);
  EHIPC(hipGetLastError());
  std::vector<float> hl(B);
  EHIPC(hipMemcpy(host_grad, grad, NP * 4, hipMemcpyDeviceToHost));
  EHIPC(hipMemcpy(hl.data(), lt, B * 4, hipMemcpyDeviceToHost));
  double L = 0.0;
  for (float v : hl) L += v;
  *loss_out = L / (double)B;
  return PSQ_OK;
}

// Train on triplets from the exact kNN within a subset of d_data (device,
// n x 50): subset = every (n / subset)-th cell; anchors = an evenly spaced pool
// of the subset, whose exact top-k_pos within the subset gives the positives;
// negatives uniform over the subset. Each step is two launches: k_mlp_step
// (sampling, forward, loss, backward, per-block partial gradients) and
// k_mlp_reduce_adam (the ordered sum and the Adam step).
int psq_encoder_train(psq_encoder* e, const float* d_data, uint64_t n, uint64_t d,
                      const psq_encoder_train_opts* o, void* stream) {
  if (!e || !d_data || !o) return fail(PSQ_E_INVALID, "NULL argument");
  if (d != 50) return fail(PSQ_E_UNSUPPORTED, "the encoder takes 50 dimensions");
  if (n < 2) return fail(PSQ_E_INVALID, "need at least 2 cells");
  if (n >= (1ull << 31)) return fail(PSQ_E_UNSUPPORTED, "n must be < 2^31");
  std::unique_lock<std::recursive_mutex> dev_lock;  // the exact kNN below uses the workspace
  ETRY(psq_host::lock_device(&dev_lock));
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t ev0, ev1, ev2;
  EHIPC(hipEventCreate(&ev0));
  EHIPC(hipEventCreate(&ev1));
  EHIPC(hipEventCreate(&ev2));
  EHIPC(hipEventRecord(ev0, s));
  const uint64_t ns = std::min<uint64_t>(n, o->subset ? o->subset : 200000);
  const uint64_t A = std::min<uint64_t>(ns, o->anchors ? o->anchors : 4096);
  const uint32_t kp = (uint32_t)std::min<uint64_t>(o->k_pos ? o->k_pos : 15, ns - 1);
  const uint32_t B = (uint32_t)(o->batch ? o->batch : 8192);
  const uint64_t steps = o->steps ? o->steps : 100;
  const float margin = o->margin > 0 ? o->margin : 0.01f;
  const float lr = o->lr > 0 ? o->lr : 3e-3f;
  if (B > (1u << 24)) return fail(PSQ_E_INVALID, "batch must be <= 2^24");
  // subset: evenly spaced cells; pool: evenly spaced subset rows
  std::vector<uint32_t> hsub(ns), hpool(A);
  for (uint64_t i = 0; i < ns; ++i) hsub[i] = (uint32_t)((i * n) / ns);
  for (uint64_t a = 0; a < A; ++a) hpool[a] = (uint32_t)((a * ns) / A);
  const uint32_t nb = blocks_for(B, MLP_STEP_T);
  const uint32_t chunk = 1024, nmb = blocks_for(ns, chunk);
  uint32_t *d_sub, *d_pool, *d_knn;
  float *d_x, *d_knd, *lt, *partial;
  double* mom;
  ETRY(ebuf(e->sub, ns, &d_sub));
  ETRY(ebuf(e->x, ns * 50, &d_x));
  ETRY(ebuf(e->pool, A, &d_pool));
  ETRY(ebuf(e->knn, A * kp, &d_knn));
  ETRY(ebuf(e->knd, A * kp, &d_knd));
  ETRY(ebuf(e->lt, B, &lt));
  ETRY(ebuf(e->partial, (size_t)nb * NP, &partial));
  ETRY(ebuf(e->mom, nmb * 2, &mom));
  EHIPC(hipMemcpyAsync(d_sub, hsub.data(), ns * 4, hipMemcpyHostToDevice, s));
  EHIPC(hipMemcpyAsync(d_pool, hpool.data(), A * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(ns * 50, 256)), dim3(256), 0, s, d_data, 50u,
                     d_sub, (uint32_t)ns, d_x);
  EHIPC(hipGetLastError());
  // standardisation from the subset's moments (fixed-order reduction)
  hipLaunchKernelGGL(k_moments, dim3(nmb), dim3(256), 0, s, d_x, (uint32_t)ns, 50u, chunk, mom);
  EHIPC(hipGetLastError());
  std::vector<double> hm(nmb * 2);
  EHIPC(hipMemcpyAsync(hm.data(), mom, nmb * 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  EHIPC(hipStreamSynchronize(s));
  double s1 = 0, s2 = 0;
  for (uint32_t b = 0; b < nmb; ++b) {
    s1 += hm[2 * b];
    s2 += hm[2 * b + 1];
  }
  const double cnt = (double)ns * 50.0;
  const double mean = s1 / cnt, var = std::max(1e-12, s2 / cnt - mean * mean);
  e->mu = (float)mean;
  e->inv_scale = (float)(1.0 / std::sqrt(var));
  // exact top-kp of every anchor within the subset (the positives)
  ETRY(psq_host::knn_exact_device_impl(d_x, ns, 50, d_pool, A, kp, d_knn, d_knd, s));
  EHIPC(hipEventRecord(ev1, s));
  // Adam
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double c1 = 1.0, c2 = 1.0;
  EHIPC(hipFuncSetAttribute((const void*)k_mlp_step, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)mlp_step_smem()));
  for (uint64_t t = 0; t < steps; ++t) {
    hipLaunchKernelGGL(k_mlp_step, dim3(nb), dim3(64 * MLP_STEP_WAVES), mlp_step_smem(), s, d_x, nullptr, B, d_pool, (uint32_t)A,
                       d_knn, kp, (uint32_t)ns, o->seed, e->steps_done + t, e->wr, e->mu,
                       e->inv_scale, margin, partial, lt);
    c1 *= b1;
    c2 *= b2;
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(blocks_for(NP, 32)), dim3(1024), 0, s, partial, nb,
                       e->w, e->wr, e->m1, e->m2, lr, b1, b2, eps, (float)(1.0 - c1),
                       (float)(1.0 - c2), nullptr);
    EHIPC(hipGetLastError());
  }
  std::vector<float> hl(B);
  EHIPC(hipMemcpyAsync(hl.data(), lt, B * 4, hipMemcpyDeviceToHost, s));
  EHIPC(hipEventRecord(ev2, s));
  EHIPC(hipStreamSynchronize(s));
  double L = 0;
  for (float v : hl) L += v;
  e->last_loss = steps > 0 ? L / B : 0.0;
  float ms = 0, ms_mine = 0;
  (void)hipEventElapsedTime(&ms, ev0, ev2);
  (void)hipEventElapsedTime(&ms_mine, ev0, ev1);
  e->train_ms = ms;
  e->mine_ms = ms_mine;
  e->steps_done += steps;
  e->trained = true;
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  (void)hipEventDestroy(ev2);
  return PSQ_OK;
}

int psq_encoder_info(const psq_encoder* e, double* last_loss, double* train_ms, uint64_t* steps) {
  if (!e) return fail(PSQ_E_INVALID, "NULL encoder");
  if (last_loss) *last_loss = e->last_loss;
  if (train_ms) *train_ms = e->train_ms;
  if (steps) *steps = e->steps_done;
  return PSQ_OK;
}

int psq_encoder_mine_ms(const psq_encoder* e, double* mine_ms) {
  if (!e || !mine_ms) return fail(PSQ_E_INVALID, "NULL argument");
  *mine_ms = e->mine_ms;
  return PSQ_OK;
}

// Codes for n cells of d_data (device, n x 50) into d_codes (device, n x 32).
int psq_encoder_encode(const psq_encoder* e, const float* d_data, uint64_t n, uint64_t d,
                       float* d_codes, void* stream) {
  if (!e || !d_data || !d_codes) return fail(PSQ_E_INVALID, "NULL argument");
  if (d != 50) return fail(PSQ_E_UNSUPPORTED, "the encoder takes 50 dimensions");
  if (n == 0) return PSQ_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n >= (1ull << 32) - 16) return fail(PSQ_E_UNSUPPORTED, "n must be < 2^32 - 16");
  hipLaunchKernelGGL(k_encode_pack, dim3(blocks_for(ENC_PACKED, 256)), dim3(256), 0, s, e->wr,
                     e->packed);
  EHIPC(hipGetLastError());
  // persistent blocks: a few per CU, each stages the operands in LDS once
  const unsigned tiles = (unsigned)((n + 15) / 16);
  const unsigned grid = std::min<unsigned>(blocks_for(tiles, ENC_WAVES), 4096u);
  hipLaunchKernelGGL(k_encode_mfma, dim3(grid), dim3(ENC_WAVES * 64), 0, s, d_data, (uint32_t)n,
                     e->packed, e->wr, e->mu, e->inv_scale, d_codes);
  EHIPC(hipGetLastError());
  return PSQ_OK;
}

}  // extern "C"
