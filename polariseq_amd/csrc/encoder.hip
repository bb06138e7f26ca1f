// C-ABI of the learned encoder: create, train on exact-kNN triplets, encode.
// No reference counterpart: the reference has no learned component. The codes
// seed NN-Descent (PSQ_INIT_LEARNED) in place of the forest's start on the
// raw embedding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/psq_knn.h"
#include "psq_kernels.hpp"

using namespace psq;

namespace psq_host {
int fail(int code, const char* fmt, ...);
int knn_exact_device_impl(const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries,
                          uint64_t nq, uint64_t k, uint32_t* d_idx, float* d_dist, hipStream_t s);
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

void psq_encoder_destroy(psq_encoder* e) {
  if (!e) return;
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
  const uint32_t R = (uint32_t)(3 * B);
  const uint32_t nb = blocks_for(R, 16);
  uint32_t* d_rows;
  float *hpre, *z, *dz, *lt, *partial, *grad;
  EHIPC(hipMalloc(&d_rows, R * 4));
  EHIPC(hipMalloc(&hpre, (size_t)R * 128 * 4));
  EHIPC(hipMalloc(&z, (size_t)R * 32 * 4));
  EHIPC(hipMalloc(&dz, (size_t)R * 32 * 4));
  EHIPC(hipMalloc(&lt, B * 4));
  EHIPC(hipMalloc(&partial, (size_t)nb * NP * 4));
  EHIPC(hipMalloc(&grad, NP * 4));
  EHIPC(hipMemcpy(d_rows, rows, R * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_forward, dim3(nb), dim3(128), 0, 0, d_data, 50u, d_rows, R, e->wr, e->mu,
                     e->inv_scale, hpre, z);
  hipLaunchKernelGGL(k_triplet_loss, dim3(blocks_for(B, 128)), dim3(128), 0, 0, z, (uint32_t)B,
                     margin, dz, lt);
  hipLaunchKernelGGL(k_mlp_backward, dim3(nb), dim3(128), 0, 0, d_data, 50u, d_rows, R, e->wr, e->mu,
                     e->inv_scale, hpre, dz, partial);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks_for(NP, 256)), dim3(256), 0, 0, partial, nb, grad);
  EHIPC(hipGetLastError());
  std::vector<float> hl(B);
  EHIPC(hipMemcpy(host_grad, grad, NP * 4, hipMemcpyDeviceToHost));
  EHIPC(hipMemcpy(hl.data(), lt, B * 4, hipMemcpyDeviceToHost));
  double L = 0.0;
  for (float v : hl) L += v;
  *loss_out = L / (double)B;
  (void)hipFree(d_rows);
  (void)hipFree(hpre);
  (void)hipFree(z);
  (void)hipFree(dz);
  (void)hipFree(lt);
  (void)hipFree(partial);
  (void)hipFree(grad);
  return PSQ_OK;
}

// Train on triplets from the exact kNN of a subset of d_data (device, n x 50):
// subset = every (n / subset)-th cell; positives from its exact top-k_pos.
int psq_encoder_train(psq_encoder* e, const float* d_data, uint64_t n, uint64_t d,
                      const psq_encoder_train_opts* o, void* stream) {
  if (!e || !d_data || !o) return fail(PSQ_E_INVALID, "NULL argument");
  if (d != 50) return fail(PSQ_E_UNSUPPORTED, "the encoder takes 50 dimensions");
  if (n < 2) return fail(PSQ_E_INVALID, "need at least 2 cells");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t ev0, ev1;
  EHIPC(hipEventCreate(&ev0));
  EHIPC(hipEventCreate(&ev1));
  EHIPC(hipEventRecord(ev0, s));
  const uint64_t ns = std::min<uint64_t>(n, o->subset ? o->subset : 200000);
  const uint32_t kp = (uint32_t)std::min<uint64_t>(o->k_pos ? o->k_pos : 15, ns - 1);
  const uint32_t B = (uint32_t)(o->batch ? o->batch : 2048);
  const uint64_t steps = o->steps ? o->steps : 1000;
  const float margin = o->margin > 0 ? o->margin : 1.0f;
  const float lr = o->lr > 0 ? o->lr : 1e-3f;
  // subset: evenly spaced cells
  std::vector<uint32_t> hsub(ns);
  for (uint64_t i = 0; i < ns; ++i) hsub[i] = (uint32_t)((i * n) / ns);
  uint32_t *d_sub, *d_knn, *d_rows;
  float *d_x, *d_knd, *hpre, *z, *dz, *lt, *partial, *grad;
  double* mom;
  const uint32_t R = 3 * B, nb = blocks_for(R, 16);
  const uint32_t chunk = 1024, nmb = blocks_for(ns, chunk);
  EHIPC(hipMalloc(&d_sub, ns * 4));
  EHIPC(hipMalloc(&d_x, ns * 50 * 4));
  EHIPC(hipMalloc(&d_knn, ns * kp * 4));
  EHIPC(hipMalloc(&d_knd, ns * kp * 4));
  EHIPC(hipMalloc(&d_rows, R * 4));
  EHIPC(hipMalloc(&hpre, (size_t)R * 128 * 4));
  EHIPC(hipMalloc(&z, (size_t)R * 32 * 4));
  EHIPC(hipMalloc(&dz, (size_t)R * 32 * 4));
  EHIPC(hipMalloc(&lt, B * 4));
  EHIPC(hipMalloc(&partial, (size_t)nb * NP * 4));
  EHIPC(hipMalloc(&grad, NP * 4));
  EHIPC(hipMalloc(&mom, nmb * 2 * sizeof(double)));
  EHIPC(hipMemcpyAsync(d_sub, hsub.data(), ns * 4, hipMemcpyHostToDevice, s));
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
  // exact kNN of the subset (positives)
  ETRY(psq_host::knn_exact_device_impl(d_x, ns, 50, nullptr, ns, kp, d_knn, d_knd, s));
  // Adam
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double c1 = 1.0, c2 = 1.0;
  for (uint64_t t = 0; t < steps; ++t) {
    hipLaunchKernelGGL(k_triplets, dim3(blocks_for(B, 256)), dim3(256), 0, s, d_knn, kp, (uint32_t)ns,
                       B, o->seed, e->steps_done + t, d_rows);
    hipLaunchKernelGGL(k_mlp_forward, dim3(nb), dim3(128), 0, s, d_x, 50u, d_rows, R, e->wr, e->mu,
                       e->inv_scale, hpre, z);
    hipLaunchKernelGGL(k_triplet_loss, dim3(blocks_for(B, 128)), dim3(128), 0, s, z, B, margin, dz, lt);
    hipLaunchKernelGGL(k_mlp_backward, dim3(nb), dim3(128), 0, s, d_x, 50u, d_rows, R, e->wr, e->mu,
                       e->inv_scale, hpre, dz, partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks_for(NP, 256)), dim3(256), 0, s, partial, nb, grad);
    c1 *= b1;
    c2 *= b2;
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(NP, 256)), dim3(256), 0, s, e->w, e->wr, grad, e->m1,
                       e->m2, lr, b1, b2, eps, (float)(1.0 - c1), (float)(1.0 - c2));
    EHIPC(hipGetLastError());
  }
  std::vector<float> hl(B);
  EHIPC(hipMemcpyAsync(hl.data(), lt, B * 4, hipMemcpyDeviceToHost, s));
  EHIPC(hipEventRecord(ev1, s));
  EHIPC(hipStreamSynchronize(s));
  double L = 0;
  for (float v : hl) L += v;
  e->last_loss = L / B;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  e->train_ms = ms;
  e->steps_done += steps;
  e->trained = true;
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  for (void* p : {(void*)d_sub, (void*)d_x, (void*)d_knn, (void*)d_knd, (void*)d_rows, (void*)hpre,
                  (void*)z, (void*)dz, (void*)lt, (void*)partial, (void*)grad, (void*)mom})
    (void)hipFree(p);
  return PSQ_OK;
}

int psq_encoder_info(const psq_encoder* e, double* last_loss, double* train_ms, uint64_t* steps) {
  if (!e) return fail(PSQ_E_INVALID, "NULL encoder");
  if (last_loss) *last_loss = e->last_loss;
  if (train_ms) *train_ms = e->train_ms;
  if (steps) *steps = e->steps_done;
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
