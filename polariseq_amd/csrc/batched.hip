// Batched exact kNN: many independent samples (uneven sizes) in one persistent
// launch per pass, with tile shapes chosen by the project's autotuner.
//
// Reference analogue: Scrublet's per-sample exact kNN over cells plus
// simulated doublets (crates/polariseq-core/src/doublets.rs:557-651 per batch,
// scheduled largest first at 1366-1374). Order and exactness are those of
// psq_knn_exact: (FP64 distance, index), self excluded, within the sample.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/psq_knn.h"
#include "psq_kernels.hpp"

using namespace psq;

namespace psq_host {
int fail(int code, const char* fmt, ...);
}
using psq_host::fail;

#define BHIPC(call)                                                                         \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(e_ == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE, "%s failed: %s", \
                  #call, hipGetErrorString(e_));                                            \
  } while (0)
#define BTRY(expr)                 \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != PSQ_OK) return rc_; \
  } while (0)

namespace {

struct Tile {
  uint32_t block;    // threads per block (four waves of 16 or 32 queries)
  uint32_t tile;     // points per LDS tile
  uint32_t seg_len;  // points per work unit (the tuned choice)
};

// Cache of tuned tiles per shape class: (dims, K, log2 of the largest sample).
std::mutex g_tune_mu;
std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Tile> g_tuned;
thread_local psq_batched_report g_report;

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <typename T>
int dalloc(DevBuf& b, size_t count, T** out) {
  hipError_t e = hipMalloc(&b.p, std::max<size_t>(count * sizeof(T), 256));
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE, "hipMalloc failed: %s",
                hipGetErrorString(e));
  *out = static_cast<T*>(b.p);
  return PSQ_OK;
}

unsigned gsize(size_t c, unsigned b) { return (unsigned)std::max<size_t>(1, (c + b - 1) / b); }

struct Plan {
  std::vector<uint32_t> units, units2, nseg;  // units of pass 1 and of pass 2
  std::vector<uint64_t> part_off;
  uint64_t partial_floats = 0;
};

// Units ordered largest sample first (as the reference schedules batches);
// qpb1 and qpb2 are the queries per block of the two passes.
void make_plan(const std::vector<uint64_t>& off, const Tile& t, uint32_t K, uint32_t qpb1, uint32_t qpb2,
               Plan& P) {
  const size_t S = off.size() - 1;
  std::vector<uint32_t> order(S);
  for (size_t s = 0; s < S; ++s) order[s] = (uint32_t)s;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return off[a + 1] - off[a] > off[b + 1] - off[b];
  });
  P.units.clear();
  P.units2.clear();
  P.nseg.assign(S, 0);
  P.part_off.assign(S, 0);
  uint64_t acc = 0;
  for (size_t s = 0; s < S; ++s) {
    const uint64_t ns = off[s + 1] - off[s];
    P.nseg[s] = (uint32_t)std::max<uint64_t>(1, (ns + t.seg_len - 1) / t.seg_len);
    P.part_off[s] = acc;
    acc += ns * P.nseg[s] * 4 * K;  // four lists per query and segment
  }
  P.partial_floats = acc;
  // within a sample, query blocks vary fastest: the blocks running at once
  // stream the same segment, which the L2 then serves
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<uint32_t>& u = pass == 0 ? P.units : P.units2;
    const uint32_t qpb = pass == 0 ? qpb1 : qpb2;
    for (uint32_t s : order) {
      const uint64_t ns = off[s + 1] - off[s];
      const uint32_t qb = (uint32_t)((ns + qpb - 1) / qpb);
      for (uint32_t g = 0; g < P.nseg[s]; ++g)
        for (uint32_t b = 0; b < qb; ++b) {
          u.push_back(s);
          u.push_back(b);
          u.push_back(g);
        }
    }
  }
}

// Persistent grid: as many blocks of `kern` as are resident on every CU at once
// (registers and LDS permitting), so each CU runs the most waves it can.
int occupancy_grid(const void* kern, unsigned block, size_t smem, unsigned* grid) {
  int dev = 0, cus = 0, per_cu = 0;
  BHIPC(hipGetDevice(&dev));
  BHIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  BHIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)block, smem));
  *grid = (unsigned)cus * (unsigned)std::max(per_cu, 1);
  return PSQ_OK;
}

template <int D, int K>
int run_pass1(const float* Xp, const float* xn, uint32_t d, uint32_t dp, const uint64_t* d_off,
              const uint32_t* d_units, uint32_t U, const Tile& t, const uint32_t* d_nseg,
              const uint64_t* d_poff, uint32_t k, uint32_t* d_counter, float* d_partial, unsigned grid, hipStream_t s) {
  const size_t smem = exm_smem(d);
  auto kern = k_bexact_pass1<D, K>;
  if (smem > 65536)
    BHIPC(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  BTRY(occupancy_grid((const void*)kern, t.block, smem, &grid));
  BHIPC(hipMemsetAsync(d_counter, 0, 4, s));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(t.block), smem, s, Xp, xn, d, dp, d_off, d_units, U, t.seg_len,
                     d_nseg, d_poff, k, d_counter, d_partial);
  BHIPC(hipGetLastError());
  return PSQ_OK;
}

template <int D>
int dispatch_pass1(uint32_t K, const float* Xp, const float* xn, uint32_t d, uint32_t dp,
                   const uint64_t* d_off, const uint32_t* d_units, uint32_t U, const Tile& t,
                   const uint32_t* d_nseg, const uint64_t* d_poff, uint32_t k, uint32_t* d_counter, float* d_partial,
                   unsigned grid, hipStream_t s) {
  if (K == 16) return run_pass1<D, 16>(Xp, xn, d, dp, d_off, d_units, U, t, d_nseg, d_poff, k, d_counter, d_partial, grid, s);
  if (K == 32) return run_pass1<D, 32>(Xp, xn, d, dp, d_off, d_units, U, t, d_nseg, d_poff, k, d_counter, d_partial, grid, s);
  return run_pass1<D, 64>(Xp, xn, d, dp, d_off, d_units, U, t, d_nseg, d_poff, k, d_counter, d_partial, grid, s);
}


}  // namespace

extern "C" int psq_knn_exact_batched_device(const float* d_data, uint64_t n_total, uint64_t d,
                                            const uint64_t* offsets, uint64_t n_samples, uint64_t k,
                                            uint32_t* d_idx, float* d_dist, void* stream) {
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  if (!d_data || !offsets || !d_idx || !d_dist) return fail(PSQ_E_INVALID, "NULL pointer argument");
  if (n_samples == 0 || n_total == 0 || d == 0) return fail(PSQ_E_INVALID, "empty input");
  if (k == 0) return fail(PSQ_E_INVALID, "k must be >= 1");
  if (k > 64) return fail(PSQ_E_UNSUPPORTED, "k > 64 is not supported by the exact GPU path");
  if (d > 256) return fail(PSQ_E_UNSUPPORTED, "n_dim > 256 is not supported on the GPU path");
  if (n_total >= (1ull << 31)) return fail(PSQ_E_UNSUPPORTED, "n must be < 2^31");
  if (offsets[0] != 0 || offsets[n_samples] != n_total)
    return fail(PSQ_E_INVALID, "offsets must start at 0 and end at n_total");
  for (uint64_t s = 0; s < n_samples; ++s)
    if (offsets[s + 1] <= offsets[s]) return fail(PSQ_E_INVALID, "sample %llu is empty", (unsigned long long)s);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t dp = (uint32_t)((d + 31) / 32 * 32);
  const uint32_t K = k <= 16 ? 16 : (k <= 32 ? 32 : 64);
  const uint32_t cap = 128;
  const uint32_t dfix = d == 50 ? 50u : (d == 30 ? 30u : 0u);
  const uint32_t qpb1 = exm_queries((int)dfix, (int)K), qpb2 = exm_queries((int)dfix, 0);
  std::vector<uint64_t> off(offsets, offsets + n_samples + 1);
  uint64_t maxs = 0;
  for (uint64_t i = 0; i < n_samples; ++i) maxs = std::max(maxs, off[i + 1] - off[i]);
  DevBuf bXp, bOff, bCnt, bCand, bT, bOvf, bNovf, bCounter, bXn, bPmax;
  float *Xp, *T, *xn;
  uint64_t* d_off;
  uint32_t *cnt, *cand, *ovf, *novf, *counter, *pmx;
  BTRY(dalloc(bXp, n_total * dp, &Xp));
  BTRY(dalloc(bXn, n_total, &xn));
  BTRY(dalloc(bPmax, 1, &pmx));
  BTRY(dalloc(bOff, n_samples + 1, &d_off));
  BTRY(dalloc(bCnt, n_total, &cnt));
  BTRY(dalloc(bCand, n_total * cap, &cand));
  BTRY(dalloc(bT, n_total, &T));
  BTRY(dalloc(bOvf, n_total, &ovf));
  BTRY(dalloc(bNovf, 1, &novf));
  BTRY(dalloc(bCounter, 1, &counter));
  hipLaunchKernelGGL(k_pad_rows, dim3(std::min<unsigned>(gsize(n_total * dp, 256), 65536)), dim3(256), 0, s,
                     d_data, (uint32_t)n_total, (uint32_t)d, dp, Xp);
  BHIPC(hipGetLastError());
  BHIPC(hipMemcpyAsync(d_off, off.data(), (n_samples + 1) * 8, hipMemcpyHostToDevice, s));
  BHIPC(hipMemsetAsync(cnt, 0, n_total * 4, s));
  BHIPC(hipMemsetAsync(novf, 0, 4, s));
  BHIPC(hipMemsetAsync(pmx, 0, 4, s));
  hipLaunchKernelGGL(k_exact_norms, dim3(gsize(n_total, 256)), dim3(256), 0, s, Xp, (uint32_t)n_total,
                     (uint32_t)d, dp, xn, pmx);
  BHIPC(hipGetLastError());

  // ── autotune (once per shape class) ──
  const uint32_t cls = (uint32_t)std::ceil(std::log2((double)std::max<uint64_t>(maxs, 2)));
  const auto key = std::make_tuple((uint32_t)d, K, cls);
  Tile best{256, EXM_P_TILE, 16384};
  bool have = false;
  double tune_ms = 0.0;
  uint32_t tried = 0;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      best = it->second;
      have = true;
    }
  }
  if (!have) {
    // probe: the largest sample's first rows as a single-sample problem
    uint32_t big = 0;
    for (uint32_t i = 0; i < n_samples; ++i)
      if (off[i + 1] - off[i] == maxs) big = i;
    const uint64_t pn = std::min<uint64_t>(maxs, 65536);
    std::vector<uint64_t> poff = {off[big], off[big] + pn};
    // express the probe as one sample starting at off[big]
    std::vector<uint64_t> probe_off = {0, pn};
    float best_ms = 1e30f;
    const Tile cands[] = {{256, EXM_P_TILE, 4096}, {256, EXM_P_TILE, 8192}, {256, EXM_P_TILE, 16384},
                          {256, EXM_P_TILE, 32768}};
    DevBuf bpo, bpu, bpn, bpp, bpart;
    uint64_t* d_poff2;
    uint32_t *d_units2, *d_nseg2;
    uint64_t* d_pofs;
    float* d_part2;
    BTRY(dalloc(bpo, 2, &d_poff2));
    BHIPC(hipMemcpyAsync(d_poff2, probe_off.data(), 16, hipMemcpyHostToDevice, s));
    hipEvent_t a, b;
    BHIPC(hipEventCreate(&a));
    BHIPC(hipEventCreate(&b));
    // the probe reads rows at Xp + off[big] * dp (and their norms)
    const float* Xprobe = Xp + off[big] * dp;
    const float* xnprobe = xn + off[big];
    size_t max_units = 0, max_part = 0;
    std::vector<Plan> plans(sizeof(cands) / sizeof(cands[0]));
    for (size_t c = 0; c < plans.size(); ++c) {
      make_plan(probe_off, cands[c], K, qpb1, qpb2, plans[c]);
      max_units = std::max(max_units, plans[c].units.size());
      max_part = std::max<size_t>(max_part, plans[c].partial_floats);
    }
    BTRY(dalloc(bpu, max_units, &d_units2));
    BTRY(dalloc(bpn, 1, &d_nseg2));
    BTRY(dalloc(bpp, 1, &d_pofs));
    BTRY(dalloc(bpart, max_part, &d_part2));
    for (size_t c = 0; c < plans.size(); ++c) {
      const Tile& t = cands[c];
      const Plan& P = plans[c];
      const uint32_t U = (uint32_t)(P.units.size() / 3);
      BHIPC(hipMemcpyAsync(d_units2, P.units.data(), P.units.size() * 4, hipMemcpyHostToDevice, s));
      BHIPC(hipMemcpyAsync(d_nseg2, P.nseg.data(), 4, hipMemcpyHostToDevice, s));
      BHIPC(hipMemcpyAsync(d_pofs, P.part_off.data(), 8, hipMemcpyHostToDevice, s));
      unsigned grid = 0;  // sized per kernel at its launch
      BHIPC(hipEventRecord(a, s));
      if (dfix == 50)
        BTRY(dispatch_pass1<50>(K, Xprobe, xnprobe, (uint32_t)d, dp, d_poff2, d_units2, U, t, d_nseg2, d_pofs, (uint32_t)k, counter, d_part2, grid, s));
      else if (dfix == 30)
        BTRY(dispatch_pass1<30>(K, Xprobe, xnprobe, (uint32_t)d, dp, d_poff2, d_units2, U, t, d_nseg2, d_pofs, (uint32_t)k, counter, d_part2, grid, s));
      else
        BTRY(dispatch_pass1<0>(K, Xprobe, xnprobe, (uint32_t)d, dp, d_poff2, d_units2, U, t, d_nseg2, d_pofs, (uint32_t)k, counter, d_part2, grid, s));
      BHIPC(hipEventRecord(b, s));
      BHIPC(hipEventSynchronize(b));
      float ms = 0;
      BHIPC(hipEventElapsedTime(&ms, a, b));
      tune_ms += ms;
      ++tried;
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tuned[key] = best;
  }

  // ── the three persistent passes over every sample ──
  Plan P;
  make_plan(off, best, K, qpb1, qpb2, P);
  const uint32_t U = (uint32_t)(P.units.size() / 3), U2 = (uint32_t)(P.units2.size() / 3);
  DevBuf bUnits, bUnits2, bNseg, bPoff, bPart;
  uint32_t *d_units, *d_units2, *d_nseg;
  uint64_t* d_poff;
  float* d_partial;
  BTRY(dalloc(bUnits, P.units.size(), &d_units));
  BTRY(dalloc(bUnits2, P.units2.size(), &d_units2));
  BTRY(dalloc(bNseg, n_samples, &d_nseg));
  BTRY(dalloc(bPoff, n_samples, &d_poff));
  BTRY(dalloc(bPart, P.partial_floats, &d_partial));
  BHIPC(hipMemcpyAsync(d_units, P.units.data(), P.units.size() * 4, hipMemcpyHostToDevice, s));
  BHIPC(hipMemcpyAsync(d_units2, P.units2.data(), P.units2.size() * 4, hipMemcpyHostToDevice, s));
  BHIPC(hipMemcpyAsync(d_nseg, P.nseg.data(), n_samples * 4, hipMemcpyHostToDevice, s));
  BHIPC(hipMemcpyAsync(d_poff, P.part_off.data(), n_samples * 8, hipMemcpyHostToDevice, s));
  unsigned grid = 0;  // sized per kernel at its launch
  hipEvent_t e0, e1, e2, e3;
  BHIPC(hipEventCreate(&e0));
  BHIPC(hipEventCreate(&e1));
  BHIPC(hipEventCreate(&e2));
  BHIPC(hipEventCreate(&e3));
  BHIPC(hipEventRecord(e0, s));
  if (dfix == 50)
    BTRY(dispatch_pass1<50>(K, Xp, xn, (uint32_t)d, dp, d_off, d_units, U, best, d_nseg, d_poff, (uint32_t)k, counter, d_partial, grid, s));
  else if (dfix == 30)
    BTRY(dispatch_pass1<30>(K, Xp, xn, (uint32_t)d, dp, d_off, d_units, U, best, d_nseg, d_poff, (uint32_t)k, counter, d_partial, grid, s));
  else
    BTRY(dispatch_pass1<0>(K, Xp, xn, (uint32_t)d, dp, d_off, d_units, U, best, d_nseg, d_poff, (uint32_t)k, counter, d_partial, grid, s));
  BHIPC(hipEventRecord(e1, s));
  hipLaunchKernelGGL(k_bexact_tau, dim3(gsize(n_total, 128)), dim3(128), 0, s, d_partial, d_off,
                     (uint32_t)n_samples, d_nseg, d_poff, n_total, K, (uint32_t)k, (uint32_t)d, xn, pmx, T);
  BHIPC(hipGetLastError());
  {
    const size_t smem = exm_smem((uint32_t)d);
    auto kern = dfix == 50 ? k_bexact_pass2<50> : (dfix == 30 ? k_bexact_pass2<30> : k_bexact_pass2<0>);
    if (smem > 65536)
      BHIPC(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    BTRY(occupancy_grid((const void*)kern, best.block, smem, &grid));
    BHIPC(hipMemsetAsync(counter, 0, 4, s));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(best.block), smem, s, Xp, xn, (uint32_t)d, dp, d_off, d_units2, U2,
                       best.seg_len, T, counter, cnt, cand, cap);
    BHIPC(hipGetLastError());
  }
  BHIPC(hipEventRecord(e2, s));
  hipLaunchKernelGGL(k_exact_rerank, dim3(gsize(n_total, 4)), dim3(256), 0, s, Xp, (uint32_t)n_total,
                     (uint32_t)d, dp, nullptr, (uint32_t)n_total, (uint32_t)k, cnt, cand, cap, d_idx,
                     d_dist, ovf, novf);
  BHIPC(hipGetLastError());
  uint32_t h_novf = 0;
  BHIPC(hipMemcpyAsync(&h_novf, novf, 4, hipMemcpyDeviceToHost, s));
  BHIPC(hipStreamSynchronize(s));
  if (h_novf > 0) {
    DevBuf bsd, bsi;
    double* sd;
    uint32_t* si;
    BTRY(dalloc(bsd, (size_t)h_novf * k, &sd));
    BTRY(dalloc(bsi, (size_t)h_novf * k, &si));
    hipLaunchKernelGGL(k_exact_fallback, dim3(gsize(h_novf, 64)), dim3(64), 0, s, Xp, (uint32_t)n_total,
                       (uint32_t)d, dp, nullptr, (uint32_t)k, T, ovf, h_novf, sd, si, d_idx, d_dist, d_off,
                       (uint32_t)n_samples, exact_fallback_scale(d));
    BHIPC(hipGetLastError());
  }
  BHIPC(hipEventRecord(e3, s));
  BHIPC(hipStreamSynchronize(s));
  float m1 = 0, m2 = 0, m3 = 0;
  (void)hipEventElapsedTime(&m1, e0, e1);
  (void)hipEventElapsedTime(&m2, e1, e2);
  (void)hipEventElapsedTime(&m3, e2, e3);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  (void)hipEventDestroy(e3);
  double pairs = 0;
  for (uint64_t i = 0; i < n_samples; ++i) {
    const double ns = (double)(off[i + 1] - off[i]);
    pairs += ns * ns;
  }
  psq_batched_report r;
  std::memset(&r, 0, sizeof(r));
  r.block = best.block;
  r.tile = best.tile;
  r.seg_len = best.seg_len;
  r.tuned_now = have ? 0 : 1;
  r.configs_tried = tried;
  r.tune_ms = tune_ms;
  r.units = U;
  r.overflow_queries = h_novf;
  r.ms_pass1 = m1;
  r.ms_pass2 = m2;
  r.ms_rerank = m3;
  r.pairs = pairs;
  r.ms_total = std::chrono::duration<double, std::milli>(clk::now() - t_start).count();
  g_report = r;
  return PSQ_OK;
}

extern "C" int psq_knn_exact_batched(const float* data, uint64_t n_total, uint64_t d,
                                     const uint64_t* offsets, uint64_t n_samples, uint64_t k,
                                     uint32_t* idx_out, float* dist_out, int32_t device) {
  if (!data || !offsets || !idx_out || !dist_out) return fail(PSQ_E_INVALID, "NULL pointer argument");
  int count = 0;
  BHIPC(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(PSQ_E_INVALID, "device %d out of range", device);
  BHIPC(hipSetDevice(device));
  DevBuf bx, bi, bd;
  float *dx, *dd;
  uint32_t* di;
  BTRY(dalloc(bx, n_total * d, &dx));
  BTRY(dalloc(bi, n_total * k, &di));
  BTRY(dalloc(bd, n_total * k, &dd));
  BHIPC(hipMemcpy(dx, data, n_total * d * 4, hipMemcpyHostToDevice));
  BTRY(psq_knn_exact_batched_device(dx, n_total, d, offsets, n_samples, k, di, dd, nullptr));
  BHIPC(hipMemcpy(idx_out, di, n_total * k * 4, hipMemcpyDeviceToHost));
  BHIPC(hipMemcpy(dist_out, dd, n_total * k * 4, hipMemcpyDeviceToHost));
  return PSQ_OK;
}

extern "C" int psq_batched_last_report(psq_batched_report* out) {
  if (!out) return fail(PSQ_E_INVALID, "out is NULL");
  *out = g_report;
  return PSQ_OK;
}
