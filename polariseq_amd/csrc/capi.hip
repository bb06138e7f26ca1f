// C-ABI of libpsq_knn.so: argument checks, device working set, the NN-Descent
// iteration loop and the exact kNN driver. Declarations and the reference
// interfaces each entry replaces are in include/psq_knn.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psq_knn.h"
#include "psq_kernels.hpp"
#include "psq_partition.hpp"

using namespace psq;

namespace psq_host {
thread_local std::string g_err;
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
int knn_exact_device_impl(const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries,
                          uint64_t nq, uint64_t k, uint32_t* d_idx, float* d_dist, hipStream_t s);

// One lock per device, held for a whole call of every entry point that uses
// the device's cached working set: calls from several threads on one device
// run one after the other instead of sharing buffers. Recursive, so an entry
// that calls another (training mines triplets with the exact kNN) keeps it.
std::mutex g_lock_mu;
std::map<int, std::recursive_mutex> g_dev_locks;
int lock_device(std::unique_lock<std::recursive_mutex>* out) {
  int dev = 0;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PSQ_E_DEVICE, "hipGetDevice failed: %s", hipGetErrorString(e));
  std::recursive_mutex* m;
  {
    std::lock_guard<std::mutex> g(g_lock_mu);
    m = &g_dev_locks[dev];
  }
  *out = std::unique_lock<std::recursive_mutex>(*m);
  return PSQ_OK;
}
}  // namespace psq_host

#define LOCK_DEVICE()                                  \
  std::unique_lock<std::recursive_mutex> dev_lock_;    \
  do {                                                 \
    int rc_lk_ = psq_host::lock_device(&dev_lock_);    \
    if (rc_lk_ != PSQ_OK) return rc_lk_;               \
  } while (0)

using psq_host::fail;
using psq_host::g_err;

namespace {

thread_local psq_nndescent_stats g_stats;
uint64_t g_budget = 0;  // 0 = no cap

#define HIPC(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      return fail(e_ == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE, "%s failed: %s", \
                  #call, hipGetErrorString(e_));                                           \
    }                                                                                      \
  } while (0)

#define TRY(expr)               \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != PSQ_OK) return rc_; \
  } while (0)

// Grow-only named device buffers, one set per device.
struct Workspace {
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Buf> bufs;
  uint64_t total = 0;
  void* pinned = nullptr;  // small pinned host area for counters

  int get(const char* name, size_t bytes, void** out) {
    Buf& b = bufs[name];
    if (b.bytes < bytes) {
      if (b.p) {
        hipError_t e = hipFree(b.p);
        if (e != hipSuccess) return fail(PSQ_E_DEVICE, "hipFree failed: %s", hipGetErrorString(e));
        total -= b.bytes;
        b.p = nullptr;
        b.bytes = 0;
      }
      size_t want = std::max<size_t>(bytes, 256);
      if (g_budget && total + want > g_budget)
        return fail(PSQ_E_NOMEM, "device budget exceeded: %s needs %zu bytes (%llu in use, cap %llu)",
                    name, want, (unsigned long long)total, (unsigned long long)g_budget);
      hipError_t e = hipMalloc(&b.p, want);
      if (e != hipSuccess) {
        b.p = nullptr;
        return fail(e == hipErrorOutOfMemory ? PSQ_E_NOMEM : PSQ_E_DEVICE,
                    "hipMalloc(%zu) for %s failed: %s", want, name, hipGetErrorString(e));
      }
      b.bytes = want;
      total += want;
    }
    *out = b.p;
    return PSQ_OK;
  }
  int host_counters(void** out) {
    if (!pinned) {
      hipError_t e = hipHostMalloc(&pinned, 4096, hipHostMallocDefault);
      if (e != hipSuccess) return fail(PSQ_E_DEVICE, "hipHostMalloc failed: %s", hipGetErrorString(e));
    }
    *out = pinned;
    return PSQ_OK;
  }
  void release() {
    for (auto& kv : bufs)
      if (kv.second.p) (void)hipFree(kv.second.p);
    bufs.clear();
    total = 0;
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
  }
};

std::mutex g_ws_mu;
std::map<int, Workspace> g_ws;

int workspace(Workspace** out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_ws_mu);
  *out = &g_ws[dev];
  return PSQ_OK;
}

template <typename T>
int wbuf(Workspace* ws, const char* name, size_t count, T** out) {
  void* p = nullptr;
  TRY(ws->get(name, count * sizeof(T), &p));
  *out = static_cast<T*>(p);
  return PSQ_OK;
}

inline unsigned grid_for(size_t count, unsigned block) {
  size_t g = (count + block - 1) / block;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(g, 1u << 30));
}

// Grid of `blocks` blocks for the kernels that number their blocks with
// block_id(): one dimension up to GRID_X_BLOCKS blocks, rows of GRID_X_BLOCKS
// (a multiple of XCDS) beyond. The environment variable PSQ_GRID_X_BLOCKS (a
// multiple of 8) narrows the rows, so the two-dimensional layout can be tested
// at small sizes.
constexpr uint32_t GRID_X_BLOCKS = 1u << 22;
dim3 block_grid(uint64_t blocks) {
  uint64_t gx = GRID_X_BLOCKS;
  if (const char* e = std::getenv("PSQ_GRID_X_BLOCKS")) {
    const unsigned long long v = std::strtoull(e, nullptr, 10);
    if (v >= XCDS && v <= GRID_X_BLOCKS && v % XCDS == 0) gx = v;
  }
  blocks = std::max<uint64_t>(blocks, 1);
  if (blocks <= gx) return dim3((unsigned)blocks);
  return dim3((unsigned)gx, (unsigned)((blocks + gx - 1) / gx));
}

int fill_u64(uint64_t* p, size_t count, uint64_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(std::min<unsigned>(grid_for(count, 256), 65536)), dim3(256),
                     0, s, p, count, v);
  HIPC(hipGetLastError());
  return PSQ_OK;
}
int fill_u32(uint32_t* p, size_t count, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(std::min<unsigned>(grid_for(count, 256), 65536)), dim3(256),
                     0, s, p, count, v);
  HIPC(hipGetLastError());
  return PSQ_OK;
}
int fill_u8(uint8_t* p, size_t count, uint8_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u8, dim3(std::min<unsigned>(grid_for(count, 256), 65536)), dim3(256),
                     0, s, p, count, v);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

// Exclusive scan of n u32 values into out[0..n], out[n] = total. Up to one
// block of values, the block's own total is out[n]: one launch.
int scan_u32(Workspace* ws, const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t s) {
  const uint32_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_local, dim3(1), dim3(SCAN_BLOCK), 0, s, in, out, n, out + n);
    HIPC(hipGetLastError());
    return PSQ_OK;
  }
  uint32_t* sums = nullptr;
  TRY(wbuf(ws, "scan_sums", std::max<uint32_t>(nb, 1), &sums));
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(SCAN_BLOCK), 0, s, in, out, n, sums);
    HIPC(hipGetLastError());
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, sums, nb, out + n);
  HIPC(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_BLOCK), 0, s, out, n, sums);
    HIPC(hipGetLastError());
  }
  return PSQ_OK;
}

uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// ── two-level partition (psq_partition.hpp) ──
struct PartBufs {
  uint32_t *stot, *sstart, *sfill, *btot, *bstart, *bfill, *pj1;
  uint64_t *pay1, *pay2;
  uint16_t* pl2;
};

int part_bufs(Workspace* ws, const std::string& pre, const PartShape& P, uint64_t m, PartBufs* B) {
  const uint64_t mm = std::max<uint64_t>(m, 1);
  TRY(wbuf(ws, (pre + "_stot").c_str(), (size_t)P.ns + 1, &B->stot));
  TRY(wbuf(ws, (pre + "_sstart").c_str(), (size_t)P.ns + 1, &B->sstart));
  TRY(wbuf(ws, (pre + "_sfill").c_str(), P.ns, &B->sfill));
  TRY(wbuf(ws, (pre + "_btot").c_str(), (size_t)P.nb + 1, &B->btot));
  TRY(wbuf(ws, (pre + "_bstart").c_str(), (size_t)P.nb + 1, &B->bstart));
  TRY(wbuf(ws, (pre + "_bfill").c_str(), P.nb, &B->bfill));
  TRY(wbuf(ws, (pre + "_pay1").c_str(), mm, &B->pay1));
  TRY(wbuf(ws, (pre + "_pj1").c_str(), mm, &B->pj1));
  TRY(wbuf(ws, (pre + "_pay2").c_str(), mm, &B->pay2));
  TRY(wbuf(ws, (pre + "_pl2").c_str(), mm, &B->pl2));
  return PSQ_OK;
}

int part_clear(const PartShape& P, PartBufs& B, hipStream_t s) {
  TRY(fill_u32(B.stot, (size_t)P.ns + 1, 0, s));
  TRY(fill_u32(B.sfill, P.ns, 0, s));
  TRY(fill_u32(B.btot, (size_t)P.nb + 1, 0, s));
  TRY(fill_u32(B.bfill, P.nb, 0, s));
  return PSQ_OK;
}

// Level 2: super-bucket regions (pay1, pj1 from level 1, sstart) -> bucket
// regions (pay2, pl2) with bstart.
int part_level2(Workspace* ws, const PartShape& P, PartBufs& B, hipStream_t s) {
  hipLaunchKernelGGL(k_part_l2_count, dim3(P.ns, PART_L2_SHARES), dim3(1024), 0, s, B.sstart, B.pj1,
                     P.sshift, P.blog, B.btot);
  HIPC(hipGetLastError());
  TRY(scan_u32(ws, B.btot, B.bstart, P.nb, s));
  const size_t l2_smem = (size_t)PART_P2_CHUNK * (sizeof(uint64_t) + sizeof(uint32_t));
  HIPC(hipFuncSetAttribute((const void*)k_part_l2_scatter,
                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)l2_smem));
  hipLaunchKernelGGL(k_part_l2_scatter, dim3(P.ns, PART_L2_SHARES), dim3(1024), l2_smem, s, B.sstart,
                     B.pay1, B.pj1, P.sshift, P.blog, B.bstart, B.bfill, B.pay2, B.pl2);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

// Reverse lists of the rows G (n rows of k) grouped by neighbour: rev, roff.
int build_reverse(Workspace* ws, const PartShape& P, PartBufs& B, const uint64_t* G, uint32_t n,
                  uint32_t k, uint64_t seed, uint64_t iter, uint32_t* roff, uint64_t* rev,
                  hipStream_t s) {
  const uint64_t nk = (uint64_t)n * k;
  TRY(part_clear(P, B, s));
  hipLaunchKernelGGL(k_rv_l1_count, dim3((unsigned)((nk + PART_C1_CHUNK - 1) / PART_C1_CHUNK)),
                     dim3(1024), 0, s, G, nk, P.sshift, P.ns, B.stot);
  HIPC(hipGetLastError());
  TRY(scan_u32(ws, B.stot, B.sstart, P.ns, s));
  hipLaunchKernelGGL(k_rv_l1_scatter, dim3((unsigned)((nk + PART_P1_CHUNK - 1) / PART_P1_CHUNK)),
                     dim3(1024), 0, s, G, nk, k, seed, iter, P.sshift, P.ns, B.sstart, B.sfill,
                     B.pay1, B.pj1);
  HIPC(hipGetLastError());
  TRY(part_level2(ws, P, B, s));
  const size_t smem = (2 * ((size_t)1 << P.blog) + 1) * sizeof(uint32_t);
  if (smem > 65536)
    HIPC(hipFuncSetAttribute((const void*)k_rv_bucket_sort,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(k_rv_bucket_sort, dim3(P.nb), dim3(1024), smem, s, B.bstart, B.pay2, B.pl2, n,
                     P.blog, P.nb, roff, rev);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

// BF16 join: a pair passes the screen when its approximate distance is below an
// endpoint's threshold times this factor (the approximation error is a few parts
// per thousand on centred vectors); passing pairs are re-ranked exactly.
constexpr float BF16_SLACK = 1.03f;
constexpr uint32_t BF16_SV_CAP = 320;  // screened-pair list of k_join_bf16
constexpr uint64_t EX_BIG_QMIN = 256;   // smallest query chunk of the large-k exact path
// BF16 form: the join's threshold is a row's entry of rank bf16_thr_rank (1 =
// its nearest), not its heap_k-th, in the first two iterations: k (the output
// size) in the second and, when rows keep at least 1.5 k entries, 0.4 k (rounded
// up) in the first; the rows still keep heap_k entries. Early rows are poor,
// and the wider thresholds mostly buy proposals that later iterations replace
// anyway (on 4.06M x 50, k = 15, heap_k = 24, the first join sends 2.7x fewer
// proposals for the same recall and iteration count). A pair skipped in the
// first join is not joined again, so rows with no spare entries (heap_k = k)
// keep the k-th entry there: without the spares the narrow first threshold
// costs recall.
uint32_t bf16_thr_rank(uint64_t it, uint64_t ko, uint64_t k) {
  if (it == 0 && 2 * k >= 3 * ko) return (uint32_t)std::max<uint64_t>(1, (2 * ko + 4) / 5);
  if (it <= 1) return (uint32_t)ko;
  return (uint32_t)k;
}

struct Resolved {
  // ko: neighbours output per row; k: entries each row keeps while iterating.
  uint64_t n, d, k, ko, m, n_iters, n_trees, leaf_size, S, stop_at, seed;
  uint32_t dp;
};

// nndescent.rs:722-752.
int resolve(uint64_t n, uint64_t d, const psq_nndescent_opts* o, Resolved* r) {
  if (!o) return fail(PSQ_E_INVALID, "opts is NULL");
  if (n == 0 || d == 0) return fail(PSQ_E_INVALID, "empty data: n_obs=%llu n_dim=%llu",
                                    (unsigned long long)n, (unsigned long long)d);
  if (o->k == 0) return fail(PSQ_E_INVALID, "k must be >= 1");
  // row keys hold a 31-bit index
  if (n >= (1ull << 31)) return fail(PSQ_E_UNSUPPORTED, "n_obs must be < 2^31");
  if (d > 256) return fail(PSQ_E_UNSUPPORTED, "n_dim > 256 is not supported on the GPU path");
  r->n = n;
  r->d = d;
  r->ko = std::max<uint64_t>(1, std::min<uint64_t>(o->k, n - 1));
  const bool bf16 = o->precision == PSQ_PRECISION_BF16_RERANK;
  uint64_t heap = o->heap_k > 0 ? o->heap_k : (bf16 ? std::min<uint64_t>(2 * r->ko, 64) : r->ko);
  r->k = std::min<uint64_t>(std::max<uint64_t>(heap, r->ko), n - 1);
  // options the reference derives from k (candidates, leaf size) follow the
  // output k; bucket slots and the stop rule follow the row size
  r->m = std::max<uint64_t>(1, o->max_candidates > 0 ? o->max_candidates
                                                      : std::min<uint64_t>(2 * r->ko, 60));
  r->n_iters = o->max_iterations > 0
                   ? o->max_iterations
                   : std::max<uint64_t>(5, (uint64_t)std::llround(std::log2((double)n)));
  r->n_trees = o->n_trees > 0
                   ? o->n_trees
                   : std::min<uint64_t>(48, 8 + (uint64_t)std::llround(std::sqrt((double)n) / 10.0));
  r->leaf_size = o->leaf_size > 0 ? o->leaf_size : std::max<uint64_t>(r->ko, 10);
  uint64_t S = 64;
  while (S < 4 * r->k) S *= 2;
  if (o->slots) {
    S = o->slots;
    if (S < 64 || (S & (S - 1)) != 0) return fail(PSQ_E_INVALID, "slots must be a power of two >= 64");
  }
  r->S = S;
  r->stop_at = (uint64_t)(o->delta * (double)n * (double)r->k);
  r->seed = o->seed;
  r->dp = round_up((uint32_t)d, 32);
  if (r->k > 256) return fail(PSQ_E_UNSUPPORTED, "k and heap_k > 256 are not supported on the GPU path");
  // the reverse lists' offsets and the graph's row offsets are 32-bit
  if (n * r->k > 0xFFFFFF00ull)
    return fail(PSQ_E_UNSUPPORTED, "n_obs x heap_k must be below 2^32 (%llu x %llu)",
                (unsigned long long)n, (unsigned long long)r->k);
  if (r->m > 64) return fail(PSQ_E_UNSUPPORTED, "max_candidates > 64 is not supported on the GPU path");
  if (r->S > 1024) return fail(PSQ_E_UNSUPPORTED, "slots > 1024 is not supported");
  if (o->precision != PSQ_PRECISION_FP32 && o->precision != PSQ_PRECISION_BF16_RERANK)
    return fail(PSQ_E_INVALID, "unknown precision %d", o->precision);
  if (o->precision == PSQ_PRECISION_BF16_RERANK && d > 64)
    return fail(PSQ_E_UNSUPPORTED, "the BF16 form supports n_dim <= 64");
  if (o->init != PSQ_INIT_RANDOM && o->init != PSQ_INIT_RP_FOREST && o->init != PSQ_INIT_LEARNED)
    return fail(PSQ_E_INVALID, "unknown init %d", o->init);
  if (o->init == PSQ_INIT_LEARNED && !o->model)
    return fail(PSQ_E_INVALID, "init=learned needs a trained encoder (model is NULL)");
  if (o->init == PSQ_INIT_LEARNED && d != 50)
    return fail(PSQ_E_UNSUPPORTED, "the learned encoder takes 50 dimensions");
  return PSQ_OK;
}

// Merges the slots into the rows; counter word 2 += entries changed. Each row
// writes its count to chg[row] (zero when it is left alone) and one pass sums
// them, instead of an atomic per changed row on the shared counter lines.
int launch_merge(const Resolved& R, uint64_t* G, uint64_t* slots, uint8_t* touched, int force,
                 uint32_t* chg, unsigned long long* changed, hipStream_t s) {
  const dim3 grid = block_grid(R.n), block(64);
  const uint32_t n = (uint32_t)R.n, k = (uint32_t)R.k, S = (uint32_t)R.S;
  if (k > 64 || S > 256) {  // rows or slots past one wave's registers: sort in LDS
    const uint32_t need = k + S;
    auto wk = need <= 512 ? k_merge_wide<512> : (need <= 1024 ? k_merge_wide<1024> : k_merge_wide<2048>);
    hipLaunchKernelGGL(wk, grid, dim3(256), 0, s, G, n, k, slots, S, touched, force, chg);
  } else if (S == 64)
    hipLaunchKernelGGL((k_merge<1, 2>), grid, block, 0, s, G, n, k, slots, S, touched, force,
                       chg);
  else if (S == 128)
    hipLaunchKernelGGL((k_merge<2, 4>), grid, block, 0, s, G, n, k, slots, S, touched, force,
                       chg);
  else if (S == 256)
    hipLaunchKernelGGL((k_merge<4, 8>), grid, block, 0, s, G, n, k, slots, S, touched, force,
                       chg);
  else
    return fail(PSQ_E_UNSUPPORTED, "slots must be 64, 128 or 256");
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_sum_u32_to, dim3(std::min<uint32_t>(grid_for(n, 256), 1024)), dim3(256), 0, s, chg,
                     n, changed, 2u);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

// LDS row stride of the join kernels: d rounded up to 4 floats, with an odd
// number of 16-byte units so that consecutive rows start on different banks.
uint32_t join_row_stride(uint64_t d) { return (uint32_t)(round_up((uint32_t)d, 4) | 4u); }

// Rows a join block holds in the launch over every cell. Fewer rows than 2m
// mean less LDS per block and more resident blocks per CU: the join waits on
// memory, and its time per cell falls with the blocks in flight. The cells with
// more candidates are deferred to launches over lists of cells: those with at
// most JOIN_MID_ROWS candidates to one with room for that many rows, the rest
// to one with room for 2m.
constexpr uint32_t JOIN_SMALL_ROWS = 32;
constexpr uint32_t JOIN_MID_ROWS = 48;

struct JoinArgs {
  const float* Xp;
  const uint32_t *newl, *oldl, *ncnt;
  const float* thr;
  uint64_t* slots;
  uint8_t* touched;
  uint32_t* flag;   // n: cells of one deferred launch
  uint32_t* pos;    // n + 1: their exclusive scan
  uint32_t* list;   // n: cells of the launch with JOIN_MID_ROWS rows, in order
  uint32_t* list2;  // n: cells of the launch with room for 2m rows, in order
  uint32_t* hcount; // pinned host words: the two lists' lengths
  hipEvent_t counted;
  unsigned long long* ctr;
  uint32_t* rec;    // n * JOIN_REC_WORDS: the cells' join records
  bool bf16;
};

size_t join_smem(bool bf16, uint32_t rows, uint32_t RS) {
  return bf16 ? (size_t)PB_CAP * 12 + (size_t)(rows + 1) * RS * 4 + (size_t)rows * 12 +
                    (size_t)BF16_SV_CAP * 4
              : (size_t)PB_CAP * 12 + (size_t)rows * (RS * 4 + 8);
}

int launch_join(Workspace* ws, const Resolved& R, const JoinArgs& a, uint64_t iter, hipStream_t s) {
  const uint32_t n = (uint32_t)R.n, m = (uint32_t)R.m, RS = join_row_stride(R.d);
  const uint32_t full = a.bf16 ? (2 * m + 15) / 16 * 16 : 2 * m;
  auto bk = R.d == 50 ? k_join_bf16<50> : k_join_bf16<0>;
  auto pk = R.d == 50 ? k_join<50> : k_join<0>;
  const void* kern = a.bf16 ? (const void*)bk : (const void*)pk;
  auto launch = [&](uint32_t rows, JoinCells jc, uint32_t grid) -> int {
    const size_t smem = join_smem(a.bf16, rows, RS);
    if (smem > 65536)
      HIPC(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    if (a.bf16)
      hipLaunchKernelGGL(bk, block_grid(grid), dim3(64), smem, s, a.Xp, n, (uint32_t)R.d, R.dp, RS, a.newl,
                         a.oldl, a.ncnt, m, rows, jc, a.thr, BF16_SLACK, a.slots, a.touched,
                         (uint32_t)R.S, R.seed, iter, a.rec);
    else
      hipLaunchKernelGGL(pk, block_grid(grid), dim3(64), smem, s, a.Xp, n, (uint32_t)R.d, R.dp, RS, a.newl,
                         a.oldl, a.ncnt, m, rows, jc, a.thr, a.slots, a.touched, (uint32_t)R.S,
                         R.seed, iter, a.rec);
    HIPC(hipGetLastError());
    return PSQ_OK;
  };
  // the records into the counters, after every launch of the pass
  auto reduce = [&]() -> int {
    hipLaunchKernelGGL(k_join_rec_reduce, dim3(std::min<uint32_t>(grid_for(n, 256), 1024)), dim3(256), 0,
                       s, a.rec, n, a.ctr);
    HIPC(hipGetLastError());
    return PSQ_OK;
  };
  if (full <= JOIN_SMALL_ROWS) {
    TRY(launch(full, JoinCells{nullptr, 0}, xcd_grid(n)));
    return reduce();
  }
  // the cells with lo < candidates <= hi, listed in cell order, and their number
  auto defer = [&](uint32_t lo, uint32_t hi, uint32_t* list, uint32_t* hcount) -> int {
    hipLaunchKernelGGL(k_join_defer_flags, dim3(grid_for(n, 256)), dim3(256), 0, s, a.ncnt, n, lo,
                       hi, a.flag);
    HIPC(hipGetLastError());
    TRY(scan_u32(ws, a.flag, a.pos, n, s));
    hipLaunchKernelGGL(k_join_defer_list, dim3(grid_for(n, 256)), dim3(256), 0, s, a.flag, a.pos, n,
                       list);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(hcount, a.pos + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    return PSQ_OK;
  };
  const uint32_t mid = std::min(full, JOIN_MID_ROWS);
  TRY(defer(JOIN_SMALL_ROWS, mid, a.list, a.hcount));
  if (mid < full) TRY(defer(mid, full, a.list2, a.hcount + 1));
  HIPC(hipEventRecord(a.counted, s));
  TRY(launch(JOIN_SMALL_ROWS, JoinCells{nullptr, 0}, xcd_grid(n)));
  // the counts are read while the first launch runs
  HIPC(hipEventSynchronize(a.counted));
  const uint32_t c1 = *a.hcount, c2 = mid < full ? a.hcount[1] : 0u;
  if (c1) TRY(launch(mid, JoinCells{a.list, c1}, xcd_grid(c1)));
  if (c2) TRY(launch(full, JoinCells{a.list2, c2}, xcd_grid(c2)));
  return reduce();
}

double ms_between(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
  return ms;
}

struct EventPool {
  std::vector<hipEvent_t> ev;
  ~EventPool() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  int make(hipEvent_t* out) {
    hipEvent_t e;
    hipError_t rc = hipEventCreate(&e);
    if (rc != hipSuccess) return fail(PSQ_E_DEVICE, "hipEventCreate failed: %s", hipGetErrorString(rc));
    ev.push_back(e);
    *out = e;
    return PSQ_OK;
  }
};

uint32_t leaf_group(uint64_t leaf_size) {
  uint32_t g = 16;
  while (g < leaf_size && g < 64) g *= 2;
  return g;
}

template <int D>
int launch_leaf_merge(uint32_t KM, uint32_t LG, const float* Xp, uint32_t d, uint32_t dp, uint32_t n,
                      const uint32_t* ids_t, const uint32_t* gs, const uint32_t* gc, uint32_t nl,
                      uint64_t* G, uint32_t k, hipStream_t s) {
  const uint32_t RS = join_row_stride(d);
  if (KM > 64) {  // rows of more than 64 entries: one block per leaf
    const size_t wsmem = (size_t)LG * (RS + 1) * sizeof(float);
    auto wk = KM == 128 ? k_leaf_merge_wide<D, 2> : k_leaf_merge_wide<D, 4>;
    if (wsmem > 65536)
      HIPC(hipFuncSetAttribute((const void*)wk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)wsmem));
    hipLaunchKernelGGL(wk, block_grid(nl), dim3(256), wsmem, s, Xp, d, dp, RS, n, ids_t, gs, gc, nl, G,
                       k, LG);
    HIPC(hipGetLastError());
    return PSQ_OK;
  }
  const unsigned bt = dp <= 128 ? 256 : 64;
  if (bt < LG) return fail(PSQ_E_UNSUPPORTED, "leaf group larger than the block");
  const size_t smem = (size_t)bt * (RS + 1) * sizeof(float);
  const unsigned grid = grid_for(nl, bt / LG);
  void (*kern)(const float*, uint32_t, uint32_t, uint32_t, uint32_t, const uint32_t*,
               const uint32_t*, const uint32_t*, uint32_t, uint64_t*, uint32_t) = nullptr;
  // rows of the common lengths get kernels with the length fixed at compile time
  if (LG == 16 && k == 15) kern = k_leaf_merge<D, 15, 16, 15>;
  else if (LG == 16 && k == 20) kern = k_leaf_merge<D, 20, 16, 20>;
  else if (LG == 16 && k == 26) kern = k_leaf_merge<D, 26, 16, 26>;
  else if (LG == 16 && k == 30) kern = k_leaf_merge<D, 30, 16, 30>;
  else if (KM == 16 && LG == 16) kern = k_leaf_merge<D, 16, 16, 0>;
  else if (KM == 32 && LG == 16) kern = k_leaf_merge<D, 32, 16, 0>;
  else if (KM == 32 && LG == 32) kern = k_leaf_merge<D, 32, 32, 0>;
  else if (KM == 64 && LG == 16) kern = k_leaf_merge<D, 64, 16, 0>;
  else if (KM == 64 && LG == 32) kern = k_leaf_merge<D, 64, 32, 0>;
  else kern = k_leaf_merge<D, 64, 64, 0>;
  if (smem > 65536) HIPC(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(kern, block_grid(grid), dim3(bt), smem, s, Xp, d, dp, RS, n, ids_t, gs, gc, nl, G, k);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

// RP forest start: trees built level by level, a batch of trees at a time (so
// every position index fits in 32 bits). A subtree of at most `cap` points is
// finished in LDS by one block (k_forest_finish). Leaves are recorded as counts
// at their start positions, compacted per tree, and the leaf pairs of each tree
// are merged into the rows, tree after tree (nndescent.rs:767-817).
// Xs (ds dims, row stride dps) is what the hyperplanes split: the embedding
// itself, or the learned codes. Leaf pairs are measured on Xp (the embedding).
// Renumbering of the points in the first tree's leaf order (perm[p] = original
// index of point p, inv its inverse, Xq the rows in the new order), made by
// build_forest before the leaf pairs so that the graph is built in the new
// numbering: neighbours get nearby indices, so rows, buckets and candidate rows
// touched together share cache.
struct Renumbering {
  uint32_t* perm = nullptr;
  uint32_t* inv = nullptr;
  float* Xq = nullptr;
};

int build_forest(Workspace* ws, const Resolved& R, const float* Xs, uint32_t ds, uint32_t dps,
                 const float* Xp, uint64_t* G, hipStream_t s, uint32_t* hbuf,
                 Renumbering* rn = nullptr) {
  const uint32_t n = (uint32_t)R.n, d = (uint32_t)R.d, dp = R.dp, k = (uint32_t)R.k;
  const uint32_t ls = (uint32_t)R.leaf_size;
  const uint64_t T = R.n_trees;
  const uint64_t TB = std::max<uint64_t>(1, std::min<uint64_t>(T, ((1ull << 31) - 2) / n));
  const uint32_t LG = leaf_group(ls);
  uint32_t KM = k <= 16 ? 16 : (k <= 32 ? 32 : (k <= 64 ? 64 : (k <= 128 ? 128 : 256)));
  if (LG > KM) KM = LG;
  // subtrees of at most `cap` points are finished by one wave each
  const uint32_t RSs = join_row_stride(ds);
  const uint32_t cap = ls < 128 ? 256u : 0u;
  const uint32_t NC = cap / (ls + 1) + 2;
  const size_t fin_smem = ((size_t)8 * cap + 2 + (size_t)13 * NC + (size_t)2 * NC * RSs) * 4;
  if (cap > 0 && fin_smem > 65536)
    HIPC(hipFuncSetAttribute((const void*)k_forest_finish,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)fin_smem));
  // the event the level loop waits on for its counts (destroyed on return)
  struct EventGuard {
    hipEvent_t e = nullptr;
    ~EventGuard() {
      if (e) (void)hipEventDestroy(e);
    }
  } evg;
  HIPC(hipEventCreateWithFlags(&evg.e, hipEventDisableTiming));
  const hipEvent_t ev_count = evg.e;
  std::vector<uint64_t> hseeds(T);
  for (uint64_t t = 0; t < T; ++t) hseeds[t] = splitmix(R.seed ^ (t + 1));
  uint64_t* tseeds;
  TRY(wbuf(ws, "f_tseeds", T, &tseeds));
  HIPC(hipMemcpyAsync(tseeds, hseeds.data(), T * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  const size_t total_cap = (size_t)TB * n;
  const size_t node_cap = (size_t)TB * (n / (ls + 1) + 1);
  // active nodes hold more than cap points (more than leaf_size without a finish)
  const size_t act_cap = std::min(node_cap, cap > 0 ? total_cap / (cap + 1) + TB : node_cap);
  const size_t tile_cap = total_cap / FT_TILE + act_cap + 1;
  uint32_t *ids, *tmp, *fl, *sl, *lcnt_pos;
  TRY(wbuf(ws, "f_ids", total_cap, &ids));
  TRY(wbuf(ws, "f_tmp", total_cap, &tmp));
  TRY(wbuf(ws, "f_fl", total_cap, &fl));
  TRY(wbuf(ws, "f_sl", total_cap + 1, &sl));
  TRY(wbuf(ws, "f_lcnt", total_cap, &lcnt_pos));
  uint32_t *ns[2], *nc[2], *nt[2], *nlv, *mode, *nact, *nsm, *aoff, *soff, *child;
  uint32_t *sm_start, *sm_count, *sm_tree, *sm_level, *toff;
  uint32_t *ntile, *tileoff, *tnode, *tcl, *tcr, *scl, *scr;
  uint64_t *flmask, *frmask;
  float *normals, *offsets, *nabs;
  TRY(wbuf(ws, "f_ns0", act_cap, &ns[0]));
  TRY(wbuf(ws, "f_ns1", act_cap, &ns[1]));
  TRY(wbuf(ws, "f_nc0", act_cap, &nc[0]));
  TRY(wbuf(ws, "f_nc1", act_cap, &nc[1]));
  TRY(wbuf(ws, "f_nt0", act_cap, &nt[0]));
  TRY(wbuf(ws, "f_nt1", act_cap, &nt[1]));
  TRY(wbuf(ws, "f_nl", act_cap, &nlv));
  TRY(wbuf(ws, "f_mode", act_cap, &mode));
  TRY(wbuf(ws, "f_nact", act_cap, &nact));
  TRY(wbuf(ws, "f_nsm", act_cap, &nsm));
  TRY(wbuf(ws, "f_aoff", act_cap + 1, &aoff));
  TRY(wbuf(ws, "f_soff", act_cap + 1, &soff));
  TRY(wbuf(ws, "f_child", 2 * act_cap, &child));
  TRY(wbuf(ws, "f_normals", act_cap * dps, &normals));
  TRY(wbuf(ws, "f_offsets", act_cap, &offsets));
  TRY(wbuf(ws, "f_nabs", act_cap, &nabs));
  TRY(wbuf(ws, "f_ntile", act_cap, &ntile));
  TRY(wbuf(ws, "f_tileoff", act_cap + 1, &tileoff));
  TRY(wbuf(ws, "f_tnode", tile_cap, &tnode));
  TRY(wbuf(ws, "f_tcl", tile_cap, &tcl));
  TRY(wbuf(ws, "f_tcr", tile_cap, &tcr));
  TRY(wbuf(ws, "f_scl", tile_cap + 1, &scl));
  TRY(wbuf(ws, "f_scr", tile_cap + 1, &scr));
  TRY(wbuf(ws, "f_flmask", tile_cap * (FT_TILE / 64), &flmask));
  TRY(wbuf(ws, "f_frmask", tile_cap * (FT_TILE / 64), &frmask));
  // FP16 copy of the split rows for the margin screen (64- or 128-byte rows)
  const uint32_t dh = ds <= 32 ? 32u : round_up(ds, 64);
  _Float16* Xh;
  TRY(wbuf(ws, "f_xh", (size_t)n * dh, &Xh));
  hipLaunchKernelGGL(k_rows_to_half, dim3(grid_for((size_t)n * dh, 256)), dim3(256), 0, s, Xs, ds,
                     dps, n, Xh, dh);
  HIPC(hipGetLastError());

  TRY(wbuf(ws, "f_sm_start", node_cap, &sm_start));
  TRY(wbuf(ws, "f_sm_count", node_cap, &sm_count));
  TRY(wbuf(ws, "f_sm_tree", node_cap, &sm_tree));
  TRY(wbuf(ws, "f_sm_level", node_cap, &sm_level));
  TRY(wbuf(ws, "f_toff", TB + 1, &toff));
  // leaves: each leaf holds at least one point, so a batch has at most TB * n
  uint32_t *gs, *gc;
  TRY(wbuf(ws, "f_gs", total_cap, &gs));
  TRY(wbuf(ws, "f_gc", total_cap, &gc));

  for (uint64_t t0 = 0; t0 < T; t0 += TB) {
    const uint32_t tb = (uint32_t)std::min<uint64_t>(TB, T - t0);
    const size_t total = (size_t)tb * n;
    hipLaunchKernelGGL(k_forest_init, dim3(grid_for(total, 256)), dim3(256), 0, s, ids, n, tb);
    HIPC(hipGetLastError());
    TRY(fill_u32(lcnt_pos, total, 0, s));
    std::vector<uint32_t> h0s(tb, 0), h0c(tb, n), h0t(tb), h0l(tb, 0);
    for (uint32_t t = 0; t < tb; ++t) h0t[t] = t;
    uint32_t A = 0, small_base = 0;
    if (n <= ls) {
      // every tree is one leaf
      for (uint32_t t = 0; t < tb; ++t)
        HIPC(hipMemcpyAsync(lcnt_pos + (size_t)t * n, &h0c[t], 4, hipMemcpyHostToDevice, s));
    } else if (n <= cap) {
      HIPC(hipMemcpyAsync(sm_start, h0s.data(), tb * 4, hipMemcpyHostToDevice, s));
      HIPC(hipMemcpyAsync(sm_count, h0c.data(), tb * 4, hipMemcpyHostToDevice, s));
      HIPC(hipMemcpyAsync(sm_tree, h0t.data(), tb * 4, hipMemcpyHostToDevice, s));
      HIPC(hipMemcpyAsync(sm_level, h0l.data(), tb * 4, hipMemcpyHostToDevice, s));
      small_base = tb;
    } else {
      A = tb;
      HIPC(hipMemcpyAsync(ns[0], h0s.data(), tb * 4, hipMemcpyHostToDevice, s));
      HIPC(hipMemcpyAsync(nc[0], h0c.data(), tb * 4, hipMemcpyHostToDevice, s));
      HIPC(hipMemcpyAsync(nt[0], h0t.data(), tb * 4, hipMemcpyHostToDevice, s));
    }
    int cur = 0;
    for (uint32_t L = 0; A > 0; ++L) {
      const int nx = cur ^ 1;
      // tiles of the active nodes
      hipLaunchKernelGGL(k_ftile_count, dim3(grid_for(A, 256)), dim3(256), 0, s, nc[cur], A, ntile);
      HIPC(hipGetLastError());
      TRY(scan_u32(ws, ntile, tileoff, A, s));
      HIPC(hipMemcpyAsync(hbuf, tileoff + A, 4, hipMemcpyDeviceToHost, s));
      HIPC(hipEventRecord(ev_count, s));
      // the planes need no tile count: they run while the host reads it
      hipLaunchKernelGGL(k_forest_planes, dim3(grid_for(A, 4)), dim3(256), 0, s, Xs, ds, dps, n,
                         ids, ns[cur], nc[cur], nt[cur], A, tseeds + t0, L, normals, offsets,
                         nabs);
      HIPC(hipGetLastError());
      HIPC(hipEventSynchronize(ev_count));
      const uint32_t NT = hbuf[0];
      if ((size_t)NT > tile_cap)
        return fail(PSQ_E_DEVICE, "forest bookkeeping overflow (tiles %u)", NT);
      hipLaunchKernelGGL(k_ftile_map, dim3(grid_for(NT, 256)), dim3(256), 0, s, tileoff, A, NT, tnode);
      HIPC(hipGetLastError());
      hipLaunchKernelGGL(k_forest_sides_tile, dim3(NT), dim3(256), 0, s, Xs, ds, dps, n, Xh, dh, ids,
                         ns[cur], nc[cur], nt[cur], tileoff, tnode, tseeds + t0, L, normals,
                         offsets, nabs, flmask, frmask, tcl, tcr);
      HIPC(hipGetLastError());
      TRY(scan_u32(ws, tcl, scl, NT, s));
      TRY(scan_u32(ws, tcr, scr, NT, s));
      hipLaunchKernelGGL(k_forest_nodes_tile, dim3(grid_for(A, 256)), dim3(256), 0, s, nc[cur], A,
                         tileoff, scl, scr, ls, cap, L, nlv, mode, nact, nsm);
      HIPC(hipGetLastError());
      TRY(scan_u32(ws, nact, aoff, A, s));
      TRY(scan_u32(ws, nsm, soff, A, s));
      HIPC(hipMemcpyAsync(hbuf, aoff + A, 4, hipMemcpyDeviceToHost, s));
      HIPC(hipMemcpyAsync(hbuf + 1, soff + A, 4, hipMemcpyDeviceToHost, s));
      HIPC(hipEventRecord(ev_count, s));
      // the points' moves need neither count (nor anything the emit writes): they
      // run while the host reads the counts
      hipLaunchKernelGGL(k_forest_scatter_tile, dim3(NT), dim3(256), 0, s, n, ids, ns[cur], nc[cur],
                         nt[cur], tileoff, tnode, nlv, mode, flmask, frmask, scl, scr, tmp);
      HIPC(hipGetLastError());
      hipLaunchKernelGGL(k_forest_copyback_tile, dim3(NT), dim3(256), 0, s, n, ns[cur], nc[cur],
                         nt[cur], tileoff, tnode, tmp, ids);
      HIPC(hipGetLastError());
      HIPC(hipEventSynchronize(ev_count));
      const uint32_t nextA = hbuf[0], newsmall = hbuf[1];
      if ((size_t)nextA > act_cap || (size_t)small_base + newsmall > node_cap)
        return fail(PSQ_E_DEVICE, "forest bookkeeping overflow (nodes %u, subtrees %u)", nextA,
                    small_base + newsmall);
      hipLaunchKernelGGL(k_forest_emit, dim3(grid_for(A, 256)), dim3(256), 0, s, ns[cur], nc[cur],
                         nt[cur], A, n, nlv, mode, aoff, soff, ls, cap, L, ns[nx], nc[nx], nt[nx],
                         child, lcnt_pos, small_base, sm_start, sm_count, sm_tree, sm_level);
      HIPC(hipGetLastError());
      cur = nx;
      A = nextA;
      small_base += newsmall;
    }
    if (small_base > 0) {
      hipLaunchKernelGGL(k_forest_finish, block_grid(small_base), dim3(64), fin_smem, s, Xs, ds,
                         dps, RSs, n, ids, sm_start, sm_count, sm_tree, sm_level, tseeds + t0, ls,
                         cap, NC, lcnt_pos, small_base);
      HIPC(hipGetLastError());
    }
    // leaf list, grouped by tree
    hipLaunchKernelGGL(k_leaf_flags, dim3(grid_for(total, 256)), dim3(256), 0, s, lcnt_pos, total, fl);
    HIPC(hipGetLastError());
    TRY(scan_u32(ws, fl, sl, (uint32_t)total, s));
    hipLaunchKernelGGL(k_leaf_compact, dim3(grid_for(total, 256)), dim3(256), 0, s, lcnt_pos, total,
                       n, sl, gs, gc);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_tree_offsets, dim3(grid_for(tb + 1, 64)), dim3(64), 0, s, sl, n, tb, toff);
    HIPC(hipGetLastError());
    std::vector<uint32_t> hoff(tb + 1, 0);
    HIPC(hipMemcpyAsync(hoff.data(), toff, (tb + 1) * 4, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (rn) {
      if (t0 == 0) {  // tree 0: the points in leaf order
        TRY(wbuf(ws, "perm", n, &rn->perm));
        TRY(wbuf(ws, "inv", n, &rn->inv));
        TRY(wbuf(ws, "Xq", (size_t)n * dp, &rn->Xq));
        HIPC(hipMemcpyAsync(rn->perm, ids, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_perm_inverse, dim3(grid_for(n, 256)), dim3(256), 0, s, rn->perm, n,
                           rn->inv);
        HIPC(hipGetLastError());
        hipLaunchKernelGGL(k_perm_rows, dim3(grid_for((size_t)n * (dp / 4), 256)), dim3(256), 0, s,
                           Xp, dp, rn->perm, n, rn->Xq);
        HIPC(hipGetLastError());
      }
      hipLaunchKernelGGL(k_renumber, dim3(grid_for(total, 256)), dim3(256), 0, s, ids, total,
                         rn->inv);
      HIPC(hipGetLastError());
    }
    const float* Xl = rn ? rn->Xq : Xp;
    // leaf pairs, tree after tree
    for (uint32_t t = 0; t < tb; ++t) {
      const uint32_t nl = hoff[t + 1] - hoff[t];
      if (nl == 0) continue;
      const uint32_t* ids_t = ids + (size_t)t * n;
      if (d == 50)
        TRY(launch_leaf_merge<50>(KM, LG, Xl, d, dp, n, ids_t, gs + hoff[t], gc + hoff[t], nl, G, k, s));
      else
        TRY(launch_leaf_merge<0>(KM, LG, Xl, d, dp, n, ids_t, gs + hoff[t], gc + hoff[t], nl, G, k, s));
    }
  }
  return PSQ_OK;
}

int nndescent_device_impl(const float* d_data, uint64_t n_obs, uint64_t n_dim,
                          const psq_nndescent_opts* opts, uint32_t* d_indptr, uint64_t indptr_len,
                          uint32_t* d_indices, float* d_dists, uint64_t cap, uint64_t* nnz_out,
                          hipStream_t s) {
  auto t0 = std::chrono::steady_clock::now();
  Resolved R;
  TRY(resolve(n_obs, n_dim, opts, &R));
  if (indptr_len != n_obs + 1)
    return fail(PSQ_E_INVALID, "indptr must have n_obs + 1 = %llu entries (got %llu)",
                (unsigned long long)(n_obs + 1), (unsigned long long)indptr_len);
  const uint64_t need = n_obs * R.ko;
  if (cap < need)
    return fail(PSQ_E_INVALID, "outputs must hold %llu row pointers and %llu neighbours",
                (unsigned long long)(n_obs + 1), (unsigned long long)need);
  if (!d_data || !d_indptr || !d_indices || !d_dists || !nnz_out)
    return fail(PSQ_E_INVALID, "NULL pointer argument");
  if (n_obs == 1) {  // the reference's one-point graph: the point itself, at its own distance
    hipLaunchKernelGGL(k_single_point, dim3(1), dim3(64), 0, s, d_data, (uint32_t)n_dim, d_indptr,
                       d_indices, d_dists);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s));
    *nnz_out = 1;
    g_stats = psq_nndescent_stats{};
    g_stats.n_obs = 1;
    g_stats.n_dim = n_dim;
    g_stats.k = 1;
    return PSQ_OK;
  }

  Workspace* ws;
  TRY(workspace(&ws));
  const uint32_t n = (uint32_t)R.n, k = (uint32_t)R.k, m = (uint32_t)R.m, S = (uint32_t)R.S;
  float* Xp;
  uint64_t *G, *slots, *rev;
  float* thr;
  uint8_t* touched;
  uint32_t *roff, *newl, *oldl, *ncnt, *rowcnt, *jflag, *jpos, *jlist, *jlist2;
  unsigned long long* ctr;
  TRY(wbuf(ws, "Xp", (size_t)n * R.dp, &Xp));
  TRY(wbuf(ws, "G", (size_t)n * k, &G));
  TRY(wbuf(ws, "slots", (size_t)n * S, &slots));
  TRY(wbuf(ws, "thr", n, &thr));
  TRY(wbuf(ws, "touched", n, &touched));
  TRY(wbuf(ws, "roff", (size_t)n + 1, &roff));
  TRY(wbuf(ws, "rev", (size_t)n * k, &rev));
  TRY(wbuf(ws, "newl", (size_t)n * m, &newl));
  TRY(wbuf(ws, "oldl", (size_t)n * m, &oldl));
  TRY(wbuf(ws, "ncnt", n, &ncnt));
  TRY(wbuf(ws, "jflag", n, &jflag));
  TRY(wbuf(ws, "jpos", (size_t)n + 1, &jpos));
  TRY(wbuf(ws, "jlist", n, &jlist));
  TRY(wbuf(ws, "jlist2", n, &jlist2));
  TRY(wbuf(ws, "rowcnt", n, &rowcnt));
  TRY(wbuf(ws, "counters", (size_t)(CTR_STRIPES + 1) * CTR_WORDS, &ctr));
  uint32_t* jrec;
  TRY(wbuf(ws, "jrec", (size_t)n * JOIN_REC_WORDS, &jrec));
  uint32_t* mchg;
  TRY(wbuf(ws, "mchg", n, &mchg));
  const PartShape RP = part_shape(n, 8);
  PartBufs RB;
  TRY(part_bufs(ws, "rv", RP, (uint64_t)n * k, &RB));
  void* hp;
  TRY(ws->host_counters(&hp));
  unsigned long long* hctr = static_cast<unsigned long long*>(hp);

  EventPool evp;
  hipEvent_t e_start, e_init, e_iter_end, e_end;
  TRY(evp.make(&e_start));
  TRY(evp.make(&e_init));
  TRY(evp.make(&e_iter_end));
  TRY(evp.make(&e_end));
  HIPC(hipEventRecord(e_start, s));

  hipLaunchKernelGGL(k_pad_rows, dim3(std::min<unsigned>(grid_for((size_t)n * R.dp, 256), 65536)),
                     dim3(256), 0, s, d_data, n, (uint32_t)R.d, R.dp, Xp);
  HIPC(hipGetLastError());
  TRY(fill_u64(G, (size_t)n * k, EMPTY_KEY, s));
  TRY(fill_u64(slots, (size_t)n * S, EMPTY_KEY, s));
  TRY(fill_u8(touched, n, 0, s));
  TRY(fill_u64(reinterpret_cast<uint64_t*>(ctr), (size_t)(CTR_STRIPES + 1) * CTR_WORDS, 0, s));
  {
    float inf = __builtin_inff();
    uint32_t bits;
    std::memcpy(&bits, &inf, 4);
    TRY(fill_u32(reinterpret_cast<uint32_t*>(thr), n, bits, s));
  }

  // ── start: RP forest (or nothing), random fill, then one forced merge sorts every row ──
  const bool bf16 = opts->precision == PSQ_PRECISION_BF16_RERANK;
  const bool forest = opts->init != PSQ_INIT_RANDOM && R.n_trees > 0;
  const bool reorder = forest && (opts->reorder == 1 || (opts->reorder == 0 && bf16));
  Renumbering rn;
  if (opts->init == PSQ_INIT_RP_FOREST && R.n_trees > 0)
    TRY(build_forest(ws, R, Xp, (uint32_t)R.d, R.dp, Xp, G, s, static_cast<uint32_t*>(hp) + 64,
                     reorder ? &rn : nullptr));
  if (opts->init == PSQ_INIT_LEARNED && R.n_trees > 0) {
    float* codes;
    TRY(wbuf(ws, "codes", (size_t)n * 32, &codes));
    TRY(psq_encoder_encode(static_cast<const psq_encoder*>(opts->model), d_data, n, R.d, codes, s));
    TRY(build_forest(ws, R, codes, 32u, 32u, Xp, G, s, static_cast<uint32_t*>(hp) + 64,
                     reorder ? &rn : nullptr));
  }
  hipEvent_t e_forest;
  TRY(evp.make(&e_forest));
  HIPC(hipEventRecord(e_forest, s));
  const uint32_t* perm = rn.perm;
  const uint32_t* inv = rn.inv;
  if (rn.Xq) Xp = rn.Xq;
  hipLaunchKernelGGL(k_random_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, Xp, n,
                     (uint32_t)R.d, R.dp, G, k, R.seed, touched);
  HIPC(hipGetLastError());
  TRY(launch_merge(R, G, slots, touched, 1, mchg, ctr, s));
  HIPC(hipEventRecord(e_init, s));

  // ── iterations ──
  psq_nndescent_stats st;
  std::memset(&st, 0, sizeof(st));
  std::vector<hipEvent_t> jb, je, ib, ie;
  uint64_t iters = 0;
  for (uint64_t it = 0; it < R.n_iters; ++it) {
    hipEvent_t i0, i1;
    TRY(evp.make(&i0));
    TRY(evp.make(&i1));
    HIPC(hipEventRecord(i0, s));
    ib.push_back(i0);
    ie.push_back(i1);
    TRY(fill_u64(reinterpret_cast<uint64_t*>(ctr), (size_t)CTR_STRIPES * CTR_WORDS, 0, s));
    TRY(build_reverse(ws, RP, RB, G, n, k, R.seed, it, roff, rev, s));
    auto sk = (k <= 64 && S <= 256) ? k_sample_select<1>
                                    : ((k <= 128 && S <= 512) ? k_sample_select<2> : k_sample_select<4>);
    hipLaunchKernelGGL(sk, block_grid(n), dim3(64), 0, s, G, n, k, m, R.seed, it, roff, rev, newl, oldl,
                       ncnt, slots, S, thr, bf16 ? bf16_thr_rank(it, R.ko, k) : k);
    HIPC(hipGetLastError());
    hipEvent_t a, b;
    TRY(evp.make(&a));
    TRY(evp.make(&b));
    HIPC(hipEventRecord(a, s));
    hipEvent_t counted;
    TRY(evp.make(&counted));
    const JoinArgs ja{Xp,     newl,  oldl,  ncnt,  thr,     slots, touched, jflag, jpos, jlist,
                      jlist2, static_cast<uint32_t*>(hp) + 960, counted,   ctr,   jrec, bf16};
    TRY(launch_join(ws, R, ja, it, s));
    HIPC(hipEventRecord(b, s));
    jb.push_back(a);
    je.push_back(b);
    TRY(launch_merge(R, G, slots, touched, 0, mchg, ctr, s));
    unsigned long long* csum = ctr + (size_t)CTR_STRIPES * CTR_WORDS;
    hipLaunchKernelGGL(k_ctr_reduce, dim3(1), dim3(64), 0, s, ctr, csum);
    HIPC(hipGetLastError());
    HIPC(hipEventRecord(i1, s));
    HIPC(hipMemcpyAsync(hctr, csum, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    const uint64_t changed = hctr[2];
    if (it < 64) {
      st.changed[it] = changed;
      st.pairs[it] = hctr[0];
      st.cand_rows[it] = hctr[1];
      st.exact_pairs[it] = bf16 ? hctr[3] : hctr[0];
      st.proposals[it] = hctr[4];
    }
    ++iters;
    if (changed <= R.stop_at) break;
  }
  HIPC(hipEventRecord(e_iter_end, s));

  // ── extraction ──
  const uint32_t ko = (uint32_t)R.ko;
  hipLaunchKernelGGL(k_row_counts, dim3(grid_for((size_t)n * ko, 256)), dim3(256), 0, s, G, n, k, ko, inv,
                     rowcnt);
  HIPC(hipGetLastError());
  TRY(scan_u32(ws, rowcnt, d_indptr, n, s));
  hipLaunchKernelGGL(k_extract, dim3(grid_for((size_t)n * ko, 256)), dim3(256), 0, s, G, n, k,
                     ko, perm, inv, d_indptr, d_indices, d_dists);
  HIPC(hipGetLastError());
  HIPC(hipEventRecord(e_end, s));
  uint32_t nnz32 = 0;
  HIPC(hipMemcpyAsync(&nnz32, d_indptr + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  *nnz_out = nnz32;

  st.n_obs = R.n;
  st.n_dim = R.d;
  st.k = R.ko;
  st.heap_k = R.k;
  st.reordered = perm ? 1 : 0;
  st.max_candidates = R.m;
  st.max_iterations = R.n_iters;
  st.n_trees = opts->init == PSQ_INIT_RANDOM ? 0 : R.n_trees;
  st.leaf_size = R.leaf_size;
  st.slots = R.S;
  st.stop_at = R.stop_at;
  st.iterations = iters;
  st.ms_init = ms_between(e_start, e_init);
  st.ms_iterations = ms_between(e_init, e_iter_end);
  st.ms_extract = ms_between(e_iter_end, e_end);
  for (size_t i = 0; i < jb.size() && i < 64; ++i) st.ms_join[i] = ms_between(jb[i], je[i]);
  st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  g_stats = st;
  // POLARISEQ_TRACE (any value): the stage times of nndescent.rs:757-1001 on
  // stderr, as HIP-event times of this call's stream
  if (std::getenv("POLARISEQ_TRACE")) {
    const double sec = 1e-3;
    std::fprintf(stderr, "[nndescent] rp-forest build + leaf pairs (%llu trees, leaf %llu): %.4fs\n",
                 (unsigned long long)st.n_trees, (unsigned long long)R.leaf_size,
                 ms_between(e_start, e_forest) * sec);
    std::fprintf(stderr, "[nndescent] random fill: %.4fs\n", ms_between(e_forest, e_init) * sec);
    for (size_t i = 0; i < ib.size() && i < jb.size(); ++i) {
      std::fprintf(stderr, "[nndescent] iter %zu: candidates: %.4fs\n", i, ms_between(ib[i], jb[i]) * sec);
      std::fprintf(stderr, "[nndescent] iter %zu: join, %llu changed: %.4fs\n", i,
                   (unsigned long long)(i < 64 ? st.changed[i] : 0), ms_between(jb[i], ie[i]) * sec);
    }
    std::fprintf(stderr, "[nndescent] extract: %.4fs\n", st.ms_extract * sec);
    std::fprintf(stderr, "[nndescent] total: %.4fs\n", st.ms_total * sec);
  }
  return PSQ_OK;
}

// ── fuzzy connectivities (crates/polariseq-core/src/umap.rs:130-139, 327-460) ──
// fuzzy_core leaves each row's kept entries at the start of its combined
// segment in "fz_comb" and the row pointers of the packed result in "fz_optr";
// fuzzy_emit packs them. Both run in one API call, so nothing else reuses the
// buffers between them.
struct FuzzyRun {
  const uint32_t* ip = nullptr;
  uint32_t n = 0;
  uint32_t* t_ptr = nullptr;
  uint64_t* comb = nullptr;
  uint32_t* optr = nullptr;
  uint64_t total = 0;
};

int fuzzy_giant_rows(Workspace* ws, FuzzyRun& R, uint32_t* olen, const uint32_t* giant, uint32_t ng,
                     hipStream_t s) {
  uint32_t* lens;
  TRY(wbuf(ws, "fz_glen", ng, &lens));
  hipLaunchKernelGGL(k_fuzzy_row_len, dim3(grid_for(ng, 256)), dim3(256), 0, s, R.ip, R.t_ptr, giant,
                     ng, lens);
  HIPC(hipGetLastError());
  std::vector<uint32_t> rows(ng), hl(ng);
  HIPC(hipMemcpyAsync(rows.data(), giant, ng * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(hl.data(), lens, ng * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  std::map<uint32_t, std::vector<uint32_t>> by_p;  // padded length -> rows
  for (uint32_t r = 0; r < ng; ++r) {
    uint32_t P = 1;
    while (P < hl[r]) P <<= 1;
    by_p[P].push_back(rows[r]);
  }
  for (auto& kv : by_p) {
    const uint32_t P = kv.first, nr = (uint32_t)kv.second.size();
    uint32_t* cls;
    uint64_t* tmp;
    TRY(wbuf(ws, "fz_gcls", nr, &cls));
    TRY(wbuf(ws, "fz_gtmp", (size_t)nr * P, &tmp));
    HIPC(hipMemcpyAsync(cls, kv.second.data(), nr * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    const size_t tot = (size_t)nr * P;
    hipLaunchKernelGGL(k_fuzzy_gather_pad, dim3(std::min<unsigned>(grid_for(tot, 256), 65536)),
                       dim3(256), 0, s, R.ip, R.t_ptr, cls, nr, P, R.comb, tmp);
    HIPC(hipGetLastError());
    const unsigned gb = std::min<unsigned>(grid_for(tot / 2, 256), 65536);
    for (uint32_t size = 2; size <= P; size <<= 1)
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        hipLaunchKernelGGL(k_fuzzy_bitonic_step, dim3(gb), dim3(256), 0, s, tmp, nr, P, size, stride);
        HIPC(hipGetLastError());
      }
    hipLaunchKernelGGL(k_fuzzy_fold_rows, dim3(std::min<uint32_t>(nr, 1024)), dim3(FZ_BLOCK), 0, s,
                       R.ip, R.t_ptr, cls, nr, P, tmp, R.comb, olen);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s));  // cls and tmp are reused by the next class
  }
  return PSQ_OK;
}

int fuzzy_core(Workspace* ws, const uint32_t* d_ip, const uint32_t* d_ix, const float* d_dx,
               uint32_t n, uint64_t m, hipStream_t s, FuzzyRun* out) {
  FuzzyRun R;
  R.ip = d_ip;
  R.n = n;
  uint32_t *h;
  TRY(ws->host_counters((void**)&h));
  // Buckets of FZ_BUCKET_ROWS rows; super-buckets of at most PART_SB_MAX (2^26
  // rows), else PART_SB_WIDE. The emit pass counts FZ_EMIT_LANES lanes a row.
  const bool wide = n > (1u << 26);
  const PartShape P = part_shape(n, FZ_BUCKET_LOG, wide ? PART_SB_WIDE : PART_SB_MAX);
  if (P.blog != FZ_BUCKET_LOG || (uint64_t)n * FZ_EMIT_LANES >= (1ull << 32))
    return fail(PSQ_E_UNSUPPORTED, "n_obs = %u exceeds the fuzzy path's limit of 2^28 rows", n);
  const uint32_t nb = P.nb;
  uint32_t *flags, *olen, *big, *giant;
  unsigned long long* tot;
  uint64_t* pay0;
  PartBufs B;
  TRY(wbuf(ws, "fz_flags", 4, &flags));  // bad, nbig, ngiant, pad
  TRY(wbuf(ws, "fz_tot", 1, &tot));
  TRY(wbuf(ws, "fz_pay0", std::max<uint64_t>(m, 1), &pay0));
  TRY(part_bufs(ws, "fz", P, m, &B));
  TRY(wbuf(ws, "fz_tptr", (size_t)n + 1, &R.t_ptr));
  TRY(wbuf(ws, "fz_comb", std::max<uint64_t>(2 * m, 1), &R.comb));
  TRY(wbuf(ws, "fz_olen", n, &olen));
  TRY(wbuf(ws, "fz_big", n, &big));
  TRY(wbuf(ws, "fz_giant", n, &giant));
  TRY(wbuf(ws, "fz_optr", (size_t)n + 1, &R.optr));
  TRY(fill_u32(flags, 4, 0, s));
  TRY(fill_u64((uint64_t*)tot, 1, 0, s));
  // Validate before anything indexes by the input: row pointers and indices.
  hipLaunchKernelGGL(k_fuzzy_check, dim3(std::min<unsigned>(grid_for(std::max<uint64_t>(m, n), 256), 65536)),
                     dim3(256), 0, s, d_ip, n, d_ix, m, flags);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(h, flags, 4, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (h[0] & 1u) return fail(PSQ_E_INVALID, "indptr must be non-decreasing");
  if (h[0] & 2u) return fail(PSQ_E_INVALID, "an index of the graph is >= n_obs (%u)", n);
  // Phase 1: membership strengths (umap.rs:343-354).
  hipLaunchKernelGGL(k_fuzzy_strengths, dim3(grid_for(n, 256)), dim3(256), 0, s, d_ip, d_dx, n, pay0);
  HIPC(hipGetLastError());
  // Phase 2: transpose by bucket (umap.rs:357-381), two partition levels.
  TRY(part_clear(P, B, s));
  hipLaunchKernelGGL(wide ? k_fz_l1_count<PART_SB_WIDE> : k_fz_l1_count<PART_SB_MAX>, dim3(std::max<unsigned>(1, (unsigned)((m + PART_C1_CHUNK - 1) / PART_C1_CHUNK))),
                     dim3(1024), 0, s, d_ix, m, P.sshift, P.ns, B.stot);
  HIPC(hipGetLastError());
  TRY(scan_u32(ws, B.stot, B.sstart, P.ns, s));
  if (m > 0) {
    hipLaunchKernelGGL(wide ? k_fz_l1_scatter<PART_SB_WIDE> : k_fz_l1_scatter<PART_SB_MAX>, dim3((unsigned)((m + PART_P1_CHUNK - 1) / PART_P1_CHUNK)),
                       dim3(1024), 0, s, pay0, d_ix, m, P.sshift, P.ns, B.sstart, B.sfill, B.pay1,
                       B.pj1);
    HIPC(hipGetLastError());
  }
  TRY(part_level2(ws, P, B, s));
  // Phase 3: per-row sort, fuzzy union and compaction (umap.rs:488-537), by row size.
  const size_t bm_smem = (size_t)FZ_BCAP * sizeof(uint64_t);
  HIPC(hipFuncSetAttribute((const void*)k_fz_bucket_merge, hipFuncAttributeMaxDynamicSharedMemorySize,
                           (int)bm_smem));
  hipLaunchKernelGGL(k_fz_bucket_merge, dim3(nb), dim3(FZ_MB), bm_smem, s, d_ip, d_ix, pay0, n, nb,
                     B.bstart, B.pay2, B.pl2, R.t_ptr, R.comb, olen, big, flags + 1);
  HIPC(hipGetLastError());
  const size_t mid_smem = (size_t)FZ_MID_CAP * sizeof(uint64_t);
  HIPC(hipFuncSetAttribute((const void*)k_fuzzy_merge_mid, hipFuncAttributeMaxDynamicSharedMemorySize,
                           (int)mid_smem));
  hipLaunchKernelGGL(k_fuzzy_merge_mid, dim3(1024), dim3(FZ_BLOCK), mid_smem, s, d_ip, R.t_ptr, R.comb,
                     olen, big, flags + 1, giant, flags + 2);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(h, flags, 12, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (h[2] > 0) TRY(fuzzy_giant_rows(ws, R, olen, giant, h[2], s));
  TRY(scan_u32(ws, olen, R.optr, n, s));
  // At most 2m entries are kept: below 2^32 the u32 running sum is exact.
  if (2 * m < (1ull << 32)) {
    HIPC(hipMemcpyAsync(h, R.optr + n, 4, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    R.total = h[0];
  } else {
    hipLaunchKernelGGL(k_fuzzy_total, dim3(std::min<unsigned>(grid_for(n, 256), 2048)), dim3(256), 0,
                       s, olen, n, tot);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(h, tot, 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    std::memcpy(&R.total, h, 8);
  }
  if (R.total >= (1ull << 32))
    return fail(PSQ_E_UNSUPPORTED, "the connectivity graph exceeds 2^32 entries");
  *out = R;
  return PSQ_OK;
}

int fuzzy_emit(const FuzzyRun& R, uint32_t* out_ix, float* out_w, hipStream_t s) {
  if (R.total == 0) return PSQ_OK;
  hipLaunchKernelGGL(k_fuzzy_emit, dim3(grid_for((size_t)R.n * FZ_EMIT_LANES, 256)), dim3(256), 0, s,
                     R.ip, R.t_ptr, R.n, R.comb, R.optr, out_ix, out_w);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

// Reads indptr[0] and indptr[n] of a device CSR; the graph is indptr[0] == 0 based.
int fuzzy_input_nnz(const uint32_t* d_ip, uint64_t n, uint64_t* m, hipStream_t s) {
  uint32_t ends[2] = {0, 0};
  HIPC(hipMemcpyAsync(&ends[0], d_ip, 4, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(&ends[1], d_ip + n, 4, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (ends[0] != 0) return fail(PSQ_E_INVALID, "indptr[0] must be 0 (got %u)", ends[0]);
  *m = ends[1];
  return PSQ_OK;
}

// neighbors.rs:107-118 and the binding's checks (crates/polariseq-python/src/lib.rs:474-501).
int neighbors_exact_check(uint64_t n_obs, uint64_t n_dim, uint64_t n_neighbors, int32_t metric,
                          uint64_t cap) {
  if (n_obs == 0 || n_dim == 0)
    return fail(PSQ_E_INVALID, "embedding has zero rows (%llu) or dims (%llu)",
                (unsigned long long)n_obs, (unsigned long long)n_dim);
  if (n_neighbors == 0) return fail(PSQ_E_INVALID, "n_neighbors must be \xe2\x89\xa5 1, got 0");
  if (metric != 0 && metric != 1) return fail(PSQ_E_INVALID, "unknown metric %d", metric);
  const uint64_t kws = std::min<uint64_t>(n_neighbors, n_obs);
  if (cap < n_obs * (kws - 1))
    return fail(PSQ_E_INVALID, "outputs must hold %llu neighbours",
                (unsigned long long)(n_obs * (kws - 1)));
  if (n_obs >= (1ull << 31)) return fail(PSQ_E_UNSUPPORTED, "n_obs must be < 2^31");
  if (kws > NBR_K_MAX)
    return fail(PSQ_E_UNSUPPORTED, "min(n_neighbors, n_obs) > %u is not supported", NBR_K_MAX);
  if (n_obs * (kws - 1) >= (1ull << 32))
    return fail(PSQ_E_UNSUPPORTED, "the graph would hold 2^32 entries or more");
  return PSQ_OK;
}

// neighbors.rs:101-241 on the device: one heap per query, run as the reference
// runs it (graph_exact.hip), then the stable sort and entry 0 dropped.
int neighbors_exact_device_impl(const float* d_x, uint64_t n_obs, uint64_t n_dim,
                                uint64_t n_neighbors, int32_t metric, uint32_t* d_indptr,
                                uint32_t* d_indices, float* d_dists, uint64_t cap,
                                uint64_t* nnz_out, hipStream_t s) {
  TRY(neighbors_exact_check(n_obs, n_dim, n_neighbors, metric, cap));
  if (!d_x || !d_indptr || !nnz_out) return fail(PSQ_E_INVALID, "NULL pointer argument");
  const uint32_t n = (uint32_t)n_obs, d = (uint32_t)n_dim;
  const uint32_t kws = (uint32_t)std::min<uint64_t>(n_neighbors, n_obs);
  const uint32_t per = kws - 1;
  hipLaunchKernelGGL(k_nbr_indptr, dim3(grid_for((size_t)n + 1, 256)), dim3(256), 0, s, d_indptr, n,
                     per);
  HIPC(hipGetLastError());
  if (per == 0) {
    HIPC(hipStreamSynchronize(s));
    *nnz_out = 0;
    return PSQ_OK;
  }
  if (!d_indices || !d_dists) return fail(PSQ_E_INVALID, "NULL pointer argument");
  Workspace* ws;
  TRY(workspace(&ws));
  float *rs = nullptr, *hd;
  uint32_t* hi;
  TRY(wbuf(ws, "nbr_hd", (size_t)n * kws, &hd));
  TRY(wbuf(ws, "nbr_hi", (size_t)n * kws, &hi));
  if (metric == 1) {
    TRY(wbuf(ws, "nbr_rs", n, &rs));
    hipLaunchKernelGGL(k_nbr_row_norms, dim3(grid_for(n, 256)), dim3(256), 0, s, d_x, n, d, rs);
    HIPC(hipGetLastError());
  }
  const uint32_t lds = kws <= NBR_LDS_K ? 1u : 0u;
  const size_t hsmem = lds ? (size_t)kws * 64 * (sizeof(float) + sizeof(uint32_t)) : 0;
  const int dfix = d == 50 ? 50 : (d == 30 ? 30 : 0);
  void (*heap)(const float*, uint32_t, uint32_t, uint32_t, const float*, float*, uint32_t*,
               uint32_t) =
      metric == 0 ? (dfix == 50 ? k_nbr_heap<50, 0> : dfix == 30 ? k_nbr_heap<30, 0> : k_nbr_heap<0, 0>)
                  : (dfix == 50 ? k_nbr_heap<50, 1> : dfix == 30 ? k_nbr_heap<30, 1> : k_nbr_heap<0, 1>);
  hipLaunchKernelGGL(heap, dim3((n + 63) / 64), dim3(64), hsmem, s, d_x, n, d, kws, rs, hd, hi, lds);
  HIPC(hipGetLastError());
  if (kws <= 64) {
    hipLaunchKernelGGL(k_nbr_sort64, block_grid(n), dim3(64), 0, s, hd, hi, n, kws, (int)metric, d_indices,
                       d_dists);
  } else {
    uint32_t P = 1;
    while (P < kws) P <<= 1;
    const size_t ssmem = (size_t)P * sizeof(uint64_t);
    if (ssmem > 65536)
      HIPC(hipFuncSetAttribute((const void*)k_nbr_sort_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)ssmem));
    hipLaunchKernelGGL(k_nbr_sort_lds, block_grid(n), dim3(256), ssmem, s, hd, hi, n, kws, P, (int)metric,
                       d_indices, d_dists);
  }
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(s));
  *nnz_out = (uint64_t)n * per;
  return PSQ_OK;
}

int set_device(int32_t device) {
  int count = 0;
  HIPC(hipGetDeviceCount(&count));
  if (count <= 0) return fail(PSQ_E_DEVICE, "no HIP device is visible");
  if (device < 0 || device >= count)
    return fail(PSQ_E_INVALID, "device %d out of range (%d visible)", device, count);
  HIPC(hipSetDevice(device));
  return PSQ_OK;
}

}  // namespace

// Exact kNN for 64 < k <= EX_KBIG_MAX (Scrublet's k_adj): pass 1 with 64 per
// segment and at least k/16 segments, a radix-select tau per query, pass 2 into
// a candidate buffer of 2k + 256, an LDS sort per query; queries in chunks that
// bound the pass-1 and candidate buffers.
int knn_exact_big(const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries,
                  uint64_t nq, uint64_t k, uint32_t* d_idx, float* d_dist, hipStream_t s) {
  Workspace* ws;
  TRY(workspace(&ws));
  const uint32_t dp = round_up((uint32_t)d, 32);
  // Pass 1 keeps the K smallest of each segment, over 128 ceil(k / 16) / K
  // segments (about 8k values in all). With many points per segment (n >= 4 K k
  // for K = 16 or 32) the shorter lists bound the k-th as well as 64 per
  // segment, with a shorter insertion network and fewer insertions; fewer
  // points keep 64 per segment (and at most one segment per 64 points) so that
  // the union of the lists still holds k values.
  const uint32_t K = n >= 64 * k ? 16u : (n >= 32 * k ? 32u : 64u);
  // every segment gives each query four lists (one per lane group of pass 1)
  uint32_t nseg = (uint32_t)((k + 15) / 16 * 128 / K / 4);
  if (K == 64) nseg = (uint32_t)std::min<uint64_t>((n + 63) / 64, nseg);
  nseg = std::max<uint32_t>(nseg, 1);
  const uint32_t seg_len = round_up((uint32_t)((n + nseg - 1) / nseg), 64);
  nseg = (uint32_t)((n + seg_len - 1) / seg_len);
  const uint32_t cap = round_up((uint32_t)(2 * k + 256), 64);
  uint32_t pmax = 1;
  while (pmax < cap) pmax <<= 1;
  // chunk of queries: pass-1 lists and candidates within about 4 GiB
  const uint32_t L = nseg * 4 * K;  // pass-1 values per query
  const uint64_t per_q = (uint64_t)L * 4 + (uint64_t)cap * 4 + 16;
  const uint32_t qc = (uint32_t)std::max<uint64_t>(
      EX_BIG_QMIN, std::min<uint64_t>(nq, (4ull << 30) / per_q / 256 * 256));
  float *Xp, *partial, *T, *xn;
  uint32_t *cnt, *cand, *ovf, *novf, *pmx, *qidx = nullptr;
  TRY(wbuf(ws, "Xp", (size_t)n * dp, &Xp));
  TRY(wbuf(ws, "ex_xn", n, &xn));
  TRY(wbuf(ws, "ex_pmax", 1, &pmx));
  TRY(wbuf(ws, "exb_partial", (size_t)qc * L, &partial));
  TRY(wbuf(ws, "ex_T", qc, &T));
  TRY(wbuf(ws, "ex_cnt", qc, &cnt));
  TRY(wbuf(ws, "exb_cand", (size_t)qc * cap, &cand));
  TRY(wbuf(ws, "ex_ovf", qc, &ovf));
  TRY(wbuf(ws, "ex_novf", 1, &novf));
  if (!d_queries) TRY(wbuf(ws, "exb_q", qc, &qidx));
  hipLaunchKernelGGL(k_pad_rows, dim3(std::min<unsigned>(grid_for((size_t)n * dp, 256), 65536)),
                     dim3(256), 0, s, d_data, (uint32_t)n, (uint32_t)d, dp, Xp);
  HIPC(hipGetLastError());
  TRY(fill_u32(pmx, 1, 0, s));
  hipLaunchKernelGGL(k_exact_norms, dim3(grid_for(n, 256)), dim3(256), 0, s, Xp, (uint32_t)n, (uint32_t)d,
                     dp, xn, pmx);
  HIPC(hipGetLastError());
  const size_t smem = exm_smem((uint32_t)d);
  const uint32_t dfix = d == 50 ? 50u : (d == 30 ? 30u : 0u);
  void (*p1)(const float*, const float*, uint32_t, uint32_t, uint32_t, const uint32_t*, uint32_t,
             uint32_t, uint32_t, uint32_t, float*) =
      K == 16   ? (dfix == 50 ? k_exact_pass1<50, 16> : (dfix == 30 ? k_exact_pass1<30, 16> : k_exact_pass1<0, 16>))
      : K == 32 ? (dfix == 50 ? k_exact_pass1<50, 32> : (dfix == 30 ? k_exact_pass1<30, 32> : k_exact_pass1<0, 32>))
                : (dfix == 50 ? k_exact_pass1<50, 64> : (dfix == 30 ? k_exact_pass1<30, 64> : k_exact_pass1<0, 64>));
  void (*p2)(const float*, const float*, uint32_t, uint32_t, uint32_t, const uint32_t*, uint32_t,
             uint32_t, const float*, uint32_t*, uint32_t*, uint32_t) =
      dfix == 50 ? k_exact_pass2<50> : (dfix == 30 ? k_exact_pass2<30> : k_exact_pass2<0>);
  if (smem > 65536) {
    HIPC(hipFuncSetAttribute((const void*)p1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    HIPC(hipFuncSetAttribute((const void*)p2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  }
  const size_t tau_smem = (size_t)L * 4;
  if (tau_smem > 65536)
    HIPC(hipFuncSetAttribute((const void*)k_exact_tau_big, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)tau_smem));
  const size_t rr_smem = (size_t)pmax * 12;
  if (rr_smem > 65536)
    HIPC(hipFuncSetAttribute((const void*)k_exact_rerank_big,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)rr_smem));
  for (uint64_t c0 = 0; c0 < nq; c0 += qc) {
    const uint32_t qn = (uint32_t)std::min<uint64_t>(qc, nq - c0);
    const uint32_t* qp = d_queries ? d_queries + c0 : qidx;
    if (!d_queries) {
      hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(qn, 256)), dim3(256), 0, s, qidx, (uint32_t)c0, qn);
      HIPC(hipGetLastError());
    }
    const uint32_t qpb1 = exm_queries((int)dfix, (int)K), qpb2 = exm_queries((int)dfix, 0);
    hipLaunchKernelGGL(p1, dim3((qn + qpb1 - 1) / qpb1, nseg), dim3(256), smem, s, Xp, xn, (uint32_t)n, (uint32_t)d, dp,
                       qp, qn, seg_len, nseg, (uint32_t)k, partial);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_exact_tau_big, dim3(qn), dim3(256), tau_smem, s, partial, qn, L, (uint32_t)k,
                       (uint32_t)d, qp, xn, pmx, T);
    HIPC(hipGetLastError());
    TRY(fill_u32(cnt, qn, 0, s));
    TRY(fill_u32(novf, 1, 0, s));
    hipLaunchKernelGGL(p2, dim3((qn + qpb2 - 1) / qpb2, nseg), dim3(256), smem, s, Xp, xn, (uint32_t)n, (uint32_t)d, dp,
                       qp, qn, seg_len, T, cnt, cand, cap);
    HIPC(hipGetLastError());
    uint32_t* io = d_idx + c0 * k;
    float* dq = d_dist + c0 * k;
    hipLaunchKernelGGL(k_exact_rerank_big, dim3(qn), dim3(256), rr_smem, s, Xp, (uint32_t)n,
                       (uint32_t)d, dp, qp, qn, (uint32_t)k, cnt, cand, cap, pmax, io, dq, ovf, novf);
    HIPC(hipGetLastError());
    uint32_t h_novf = 0;
    HIPC(hipMemcpyAsync(&h_novf, novf, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (h_novf > 0) {
      double* sd;
      uint32_t* si;
      TRY(wbuf(ws, "ex_fb_d", (size_t)h_novf * k, &sd));
      TRY(wbuf(ws, "ex_fb_i", (size_t)h_novf * k, &si));
      hipLaunchKernelGGL(k_exact_fallback, dim3(grid_for(h_novf, 64)), dim3(64), 0, s, Xp,
                         (uint32_t)n, (uint32_t)d, dp, qp, (uint32_t)k, T, ovf, h_novf, sd, si, io, dq,
                         nullptr, 0u, exact_fallback_scale(d));
      HIPC(hipGetLastError());
      HIPC(hipStreamSynchronize(s));
    }
  }
  return PSQ_OK;
}

// Device buffers of one exact-kNN call: the library's grow-only workspace, or
// a fixed arena carved front to back at 256-byte alignment (the entry point
// that allocates nothing).
struct BufSrc {
  Workspace* ws = nullptr;
  uint8_t* base = nullptr;
  size_t cap = 0, used = 0;
  static size_t span(size_t count, size_t elem) {
    return (std::max<size_t>(count, 1) * elem + 255) & ~(size_t)255;
  }
  template <typename T>
  int take(const char* name, size_t count, T** out) {
    if (ws) return wbuf(ws, name, count, out);
    const size_t b = span(count, sizeof(T));
    if (used + b > cap)
      return fail(PSQ_E_INVALID, "scratch of %zu bytes cannot hold %s (%zu more bytes)", cap, name, b);
    *out = reinterpret_cast<T*>(base + used);
    used += b;
    return PSQ_OK;
  }
};

constexpr uint32_t EX_CAND_CAP = 128;  // pass-2 candidates per query (k <= 64 path)

// Segments of pass 1 for the k <= 64 path: about 2048 blocks in all, at least
// 64 points per segment. Any count gives the same result (each segment keeps
// K >= k values, so the k-th of the union bounds the true k-th from above).
uint32_t exact_small_nseg(uint64_t n, uint64_t nq) {
  const uint32_t qblocks = (uint32_t)((nq + 63) / 64);  // 64-query groups
  return std::max<uint32_t>(1, (uint32_t)std::min<uint64_t>((n + 63) / 64, (2048 + qblocks - 1) / qblocks));
}
uint32_t exact_small_K(uint64_t k) { return k <= 16 ? 16 : (k <= 32 ? 32 : 64); }

// Bytes BufSrc hands out to knn_exact_small with nseg segments, the
// fallback's lists reserved for every query.
size_t exact_small_bytes(uint64_t n, uint64_t d, uint64_t nq, uint64_t k, uint32_t nseg) {
  const size_t dp = round_up((uint32_t)d, 32);
  return BufSrc::span(n * dp, 4) + BufSrc::span(n, 4) + BufSrc::span(1, 4) +
         BufSrc::span(nq * nseg * 4 * exact_small_K(k), 4) +
         3 * BufSrc::span(nq, 4) + BufSrc::span(nq * EX_CAND_CAP, 4) + BufSrc::span(1, 4) +
         BufSrc::span(nq * k, 8) + BufSrc::span(nq * k, 4);
}

// The k <= 64 path: pass 1 (per-segment top K in FP32), tau, pass 2 (every
// point under the threshold), FP64 re-rank, and the fallback for overflowed
// queries.
int knn_exact_small(BufSrc& B, const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries,
                    uint64_t nq, uint64_t k, uint32_t nseg_in, uint32_t* d_idx, float* d_dist,
                    hipStream_t s) {
  const uint32_t dp = round_up((uint32_t)d, 32);
  const uint32_t K = exact_small_K(k);
  const uint32_t cap = EX_CAND_CAP;
  uint32_t nseg = std::max<uint32_t>(1, nseg_in);
  uint32_t seg_len = round_up((uint32_t)((n + nseg - 1) / nseg), 64);
  nseg = (uint32_t)((n + seg_len - 1) / seg_len);
  float *Xp, *partial, *T, *xn;
  uint32_t *cnt, *cand, *ovf, *novf, *pmx;
  TRY(B.take("Xp", (size_t)n * dp, &Xp));
  TRY(B.take("ex_xn", n, &xn));
  TRY(B.take("ex_pmax", 1, &pmx));
  TRY(B.take("ex_partial", (size_t)nq * nseg * 4 * K, &partial));
  TRY(B.take("ex_T", nq, &T));
  TRY(B.take("ex_cnt", nq, &cnt));
  TRY(B.take("ex_cand", (size_t)nq * cap, &cand));
  TRY(B.take("ex_ovf", nq, &ovf));
  TRY(B.take("ex_novf", 1, &novf));
  hipLaunchKernelGGL(k_pad_rows, dim3(std::min<unsigned>(grid_for((size_t)n * dp, 256), 65536)),
                     dim3(256), 0, s, d_data, (uint32_t)n, (uint32_t)d, dp, Xp);
  HIPC(hipGetLastError());
  TRY(fill_u32(cnt, nq, 0, s));
  TRY(fill_u32(novf, 1, 0, s));
  TRY(fill_u32(pmx, 1, 0, s));
  hipLaunchKernelGGL(k_exact_norms, dim3(grid_for(n, 256)), dim3(256), 0, s, Xp, (uint32_t)n, (uint32_t)d,
                     dp, xn, pmx);
  HIPC(hipGetLastError());
  const size_t smem = exm_smem((uint32_t)d);
  const uint32_t dfix = d == 50 ? 50u : (d == 30 ? 30u : 0u);
  const uint32_t qpb1 = exm_queries((int)dfix, (int)K), qpb2 = exm_queries((int)dfix, 0);
  const dim3 g1((uint32_t)((nq + qpb1 - 1) / qpb1), nseg), g2((uint32_t)((nq + qpb2 - 1) / qpb2), nseg);
  void (*p1)(const float*, const float*, uint32_t, uint32_t, uint32_t, const uint32_t*, uint32_t,
             uint32_t, uint32_t, uint32_t, float*);
  if (K == 16) p1 = dfix == 50 ? k_exact_pass1<50, 16> : (dfix == 30 ? k_exact_pass1<30, 16> : k_exact_pass1<0, 16>);
  else if (K == 32) p1 = dfix == 50 ? k_exact_pass1<50, 32> : (dfix == 30 ? k_exact_pass1<30, 32> : k_exact_pass1<0, 32>);
  else p1 = dfix == 50 ? k_exact_pass1<50, 64> : (dfix == 30 ? k_exact_pass1<30, 64> : k_exact_pass1<0, 64>);
  if (smem > 65536) HIPC(hipFuncSetAttribute((const void*)p1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(p1, g1, dim3(256), smem, s, Xp, xn, (uint32_t)n, (uint32_t)d, dp, d_queries,
                     (uint32_t)nq, seg_len, nseg, (uint32_t)k, partial);
  HIPC(hipGetLastError());
  // tau: a thread per query over many queries; a block per query (radix select)
  // when the queries are too few to fill the GPU and their lists fit LDS
  const uint32_t L = nseg * 4 * K;
  if (nq <= 65536 && L <= 12288)
    hipLaunchKernelGGL(k_exact_tau_big, dim3((uint32_t)nq), dim3(256), (size_t)L * 4, s, partial, (uint32_t)nq,
                       L, (uint32_t)k, (uint32_t)d, d_queries, xn, pmx, T);
  else
    hipLaunchKernelGGL(k_exact_tau, dim3(grid_for(nq, 128)), dim3(128), 0, s, partial, (uint32_t)nq, L,
                       (uint32_t)k, (uint32_t)d, d_queries, xn, pmx, T);
  HIPC(hipGetLastError());
  void (*p2)(const float*, const float*, uint32_t, uint32_t, uint32_t, const uint32_t*, uint32_t,
             uint32_t, const float*, uint32_t*, uint32_t*, uint32_t) =
      dfix == 50 ? k_exact_pass2<50> : (dfix == 30 ? k_exact_pass2<30> : k_exact_pass2<0>);
  if (smem > 65536) HIPC(hipFuncSetAttribute((const void*)p2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(p2, g2, dim3(256), smem, s, Xp, xn, (uint32_t)n, (uint32_t)d, dp, d_queries,
                     (uint32_t)nq, seg_len, T, cnt, cand, cap);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_exact_rerank, dim3(grid_for(nq, 4)), dim3(256), 0, s, Xp, (uint32_t)n,
                     (uint32_t)d, dp, d_queries, (uint32_t)nq, (uint32_t)k, cnt, cand, cap, d_idx,
                     d_dist, ovf, novf);
  HIPC(hipGetLastError());
  uint32_t h_novf = 0;
  HIPC(hipMemcpyAsync(&h_novf, novf, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (h_novf > 0) {
    double* sd;
    uint32_t* si;
    TRY(B.take("ex_fb_d", (size_t)h_novf * k, &sd));
    TRY(B.take("ex_fb_i", (size_t)h_novf * k, &si));
    hipLaunchKernelGGL(k_exact_fallback, dim3(grid_for(h_novf, 64)), dim3(64), 0, s, Xp,
                       (uint32_t)n, (uint32_t)d, dp, d_queries, (uint32_t)k, T, ovf, h_novf, sd, si,
                       d_idx, d_dist, nullptr, 0u, exact_fallback_scale(d));
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s));
  }
  return PSQ_OK;
}

// Arguments of psq_knn_exact_device and its no-allocation variant.
int exact_check(const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries, uint64_t nq,
                uint64_t k, const uint32_t* d_idx, const float* d_dist) {
  if (n == 0 || d == 0) return fail(PSQ_E_INVALID, "empty data: n=%llu d=%llu",
                                    (unsigned long long)n, (unsigned long long)d);
  if (k == 0) return fail(PSQ_E_INVALID, "k must be >= 1");
  if (!d_queries && nq != n) return fail(PSQ_E_INVALID, "queries is NULL but n_q != n");
  if (k > EX_KBIG_MAX)
    return fail(PSQ_E_UNSUPPORTED, "k > %u is not supported by the exact GPU path", EX_KBIG_MAX);
  if (d > 256) return fail(PSQ_E_UNSUPPORTED, "n_dim > 256 is not supported on the GPU path");
  if (n >= (1ull << 31) || nq >= (1ull << 31)) return fail(PSQ_E_UNSUPPORTED, "n must be < 2^31");
  if (!d_data || !d_idx || !d_dist) return fail(PSQ_E_INVALID, "NULL pointer argument");
  return PSQ_OK;
}

int psq_host::knn_exact_device_impl(const float* d_data, uint64_t n, uint64_t d,
                                    const uint32_t* d_queries, uint64_t nq, uint64_t k,
                                    uint32_t* d_idx, float* d_dist, hipStream_t s) {
  TRY(exact_check(d_data, n, d, d_queries, nq, k, d_idx, d_dist));
  if (nq == 0) return PSQ_OK;
  if (k > 64) return knn_exact_big(d_data, n, d, d_queries, nq, k, d_idx, d_dist, s);
  BufSrc B;
  TRY(workspace(&B.ws));
  return knn_exact_small(B, d_data, n, d, d_queries, nq, k, exact_small_nseg(n, nq), d_idx, d_dist,
                         s);
}

// ── exact kNN without allocation ──
// The scratch of psq_knn_exact_device_noalloc: one array per device, held from
// the library's load until the process ends. Host code never names it beyond
// hipGetSymbolAddress: its device address, looked up once per device on the
// first call there, is what the kernels receive.
static __device__ __align__(256) uint8_t scratch[PSQ_NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
std::map<int, uint8_t*> g_scratch_ptr;

int scratch_ptr(uint8_t** out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PSQ_E_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    it = g_scratch_ptr.emplace(dev, static_cast<uint8_t*>(p)).first;
  }
  *out = it->second;
  return PSQ_OK;
}

using psq_host::knn_exact_device_impl;

extern "C" {

void psq_nndescent_opts_default(psq_nndescent_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->k = 15;
  o->delta = 0.001;
  o->precision = PSQ_PRECISION_FP32;
  o->init = PSQ_INIT_RP_FOREST;
}

uint64_t psq_effective_k(uint64_t n_obs, uint64_t k) {
  uint64_t nm1 = n_obs > 0 ? n_obs - 1 : 0;
  uint64_t e = k < nm1 ? k : nm1;
  return e > 0 ? e : 1;
}

const char* psq_last_error(void) { return g_err.c_str(); }

int psq_last_stats(psq_nndescent_stats* out) {
  if (!out) return fail(PSQ_E_INVALID, "out is NULL");
  *out = g_stats;
  return PSQ_OK;
}

void psq_set_device_budget(uint64_t bytes) { g_budget = bytes; }

int psq_release_workspace(void) {
  LOCK_DEVICE();
  Workspace* ws;
  TRY(workspace(&ws));
  std::lock_guard<std::mutex> g(g_ws_mu);
  ws->release();
  return PSQ_OK;
}

int psq_nndescent_device(const float* d_data, uint64_t n_obs, uint64_t n_dim,
                         const psq_nndescent_opts* opts, uint32_t* d_indptr, uint64_t indptr_len,
                         uint32_t* d_indices, float* d_dists, uint64_t cap, uint64_t* nnz_out,
                         void* stream) {
  g_err.clear();
  LOCK_DEVICE();
  return nndescent_device_impl(d_data, n_obs, n_dim, opts, d_indptr, indptr_len, d_indices, d_dists,
                               cap, nnz_out, static_cast<hipStream_t>(stream));
}

int psq_nndescent_into(const float* data, uint64_t n_obs, uint64_t n_dim,
                       const psq_nndescent_opts* opts, uint32_t* indptr, uint64_t indptr_len,
                       uint32_t* indices, float* dists, uint64_t cap, uint64_t* nnz_out) {
  g_err.clear();
  Resolved R;
  TRY(resolve(n_obs, n_dim, opts, &R));
  if (!data || !indptr || !indices || !dists || !nnz_out)
    return fail(PSQ_E_INVALID, "NULL pointer argument");
  if (indptr_len != n_obs + 1)
    return fail(PSQ_E_INVALID, "indptr must have n_obs + 1 = %llu entries (got %llu)",
                (unsigned long long)(n_obs + 1), (unsigned long long)indptr_len);
  if (cap < n_obs * R.ko)
    return fail(PSQ_E_INVALID, "outputs must hold %llu row pointers and %llu neighbours",
                (unsigned long long)(n_obs + 1), (unsigned long long)(n_obs * R.ko));
  TRY(set_device(opts->device));
  LOCK_DEVICE();
  Workspace* ws;
  TRY(workspace(&ws));
  float* d_data;
  uint32_t *d_ip, *d_ix;
  float* d_dx;
  TRY(wbuf(ws, "io_data", (size_t)n_obs * n_dim, &d_data));
  TRY(wbuf(ws, "io_indptr", (size_t)n_obs + 1, &d_ip));
  TRY(wbuf(ws, "io_indices", (size_t)n_obs * R.ko, &d_ix));
  TRY(wbuf(ws, "io_dists", (size_t)n_obs * R.ko, &d_dx));
  HIPC(hipMemcpy(d_data, data, (size_t)n_obs * n_dim * sizeof(float), hipMemcpyHostToDevice));
  uint64_t nnz = 0;
  TRY(nndescent_device_impl(d_data, n_obs, n_dim, opts, d_ip, n_obs + 1, d_ix, d_dx,
                            n_obs * R.ko, &nnz, nullptr));
  HIPC(hipMemcpy(indptr, d_ip, (size_t)(n_obs + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(indices, d_ix, (size_t)nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(dists, d_dx, (size_t)nnz * sizeof(float), hipMemcpyDeviceToHost));
  *nnz_out = nnz;
  return PSQ_OK;
}

int psq_knn_exact_device(const float* d_data, uint64_t n, uint64_t d, const uint32_t* d_queries,
                         uint64_t n_q, uint64_t k, uint32_t* d_idx_out, float* d_dist_out,
                         void* stream) {
  g_err.clear();
  LOCK_DEVICE();
  return knn_exact_device_impl(d_data, n, d, d_queries, n_q, k, d_idx_out, d_dist_out,
                               static_cast<hipStream_t>(stream));
}

int psq_knn_exact_device_noalloc(const float* d_data, uint64_t n, uint64_t d,
                                 const uint32_t* d_queries, uint64_t n_q, uint64_t k,
                                 uint32_t* d_idx_out, float* d_dist_out, void* stream) {
  g_err.clear();
  TRY(exact_check(d_data, n, d, d_queries, n_q, k, d_idx_out, d_dist_out));
  if (k > 64) return fail(PSQ_E_UNSUPPORTED, "the allocation-free exact kNN takes k <= 64");
  if (n_q == 0) return PSQ_OK;
  // the standard segment count, halved until the buffers fit the scratch
  uint32_t nseg = exact_small_nseg(n, n_q);
  while (nseg > 1 && exact_small_bytes(n, d, n_q, k, nseg) > PSQ_NOALLOC_SCRATCH_BYTES) nseg /= 2;
  const size_t need = exact_small_bytes(n, d, n_q, k, nseg);
  if (need > PSQ_NOALLOC_SCRATCH_BYTES)
    return fail(PSQ_E_INVALID,
                "n = %llu, d = %llu, n_q = %llu, k = %llu need %zu bytes of scratch; the "
                "allocation-free entry has %llu",
                (unsigned long long)n, (unsigned long long)d, (unsigned long long)n_q,
                (unsigned long long)k, need, (unsigned long long)PSQ_NOALLOC_SCRATCH_BYTES);
  BufSrc B;
  TRY(scratch_ptr(&B.base));
  B.cap = PSQ_NOALLOC_SCRATCH_BYTES;
  return knn_exact_small(B, d_data, n, d, d_queries, n_q, k, nseg, d_idx_out, d_dist_out,
                         static_cast<hipStream_t>(stream));
}

int psq_knn_exact(const float* data, uint64_t n, uint64_t d, const uint32_t* queries, uint64_t n_q,
                  uint64_t k, uint32_t* idx_out, float* dist_out, int32_t device) {
  g_err.clear();
  if (n == 0 || d == 0) return fail(PSQ_E_INVALID, "empty data: n=%llu d=%llu",
                                    (unsigned long long)n, (unsigned long long)d);
  if (k == 0) return fail(PSQ_E_INVALID, "k must be >= 1");
  if (!queries && n_q != n) return fail(PSQ_E_INVALID, "queries is NULL but n_q != n");
  if (!data || !idx_out || !dist_out) return fail(PSQ_E_INVALID, "NULL pointer argument");
  if (queries)
    for (uint64_t i = 0; i < n_q; ++i)
      if (queries[i] >= n) return fail(PSQ_E_INVALID, "query %llu is out of range", (unsigned long long)i);
  TRY(set_device(device));
  LOCK_DEVICE();
  Workspace* ws;
  TRY(workspace(&ws));
  float* d_data;
  uint32_t *d_q = nullptr, *d_idx;
  float* d_dist;
  TRY(wbuf(ws, "io_data", (size_t)n * d, &d_data));
  if (queries) TRY(wbuf(ws, "io_queries", (size_t)n_q, &d_q));
  TRY(wbuf(ws, "io_indices", (size_t)std::max<uint64_t>(n_q * k, 1), &d_idx));
  TRY(wbuf(ws, "io_dists", (size_t)std::max<uint64_t>(n_q * k, 1), &d_dist));
  HIPC(hipMemcpy(d_data, data, (size_t)n * d * sizeof(float), hipMemcpyHostToDevice));
  if (queries) HIPC(hipMemcpy(d_q, queries, (size_t)n_q * sizeof(uint32_t), hipMemcpyHostToDevice));
  TRY(knn_exact_device_impl(d_data, n, d, d_q, n_q, k, d_idx, d_dist, nullptr));
  HIPC(hipMemcpy(idx_out, d_idx, (size_t)n_q * k * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(dist_out, d_dist, (size_t)n_q * k * sizeof(float), hipMemcpyDeviceToHost));
  return PSQ_OK;
}

// pp.neighbors' cosine metric on the approximate path (python/polariseq/__init__.py:2540-2552).
int psq_unit_rows_device(const float* d_x, uint64_t n, uint64_t d, float* d_out, void* stream) {
  if (n == 0) return PSQ_OK;
  if (!d_x || !d_out) return fail(PSQ_E_INVALID, "NULL pointer argument");
  if (d == 0) return fail(PSQ_E_INVALID, "n_dim must be >= 1");
  if (d > ROWS_D_MAX) return fail(PSQ_E_UNSUPPORTED, "n_dim > %u is not supported", ROWS_D_MAX);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_unit_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_x, n,
                     (uint32_t)d, d_out);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

int psq_cosine_distances_device(float* d_dist, uint64_t count, void* stream) {
  if (count == 0) return PSQ_OK;
  if (!d_dist) return fail(PSQ_E_INVALID, "NULL pointer argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_half_square, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, d_dist,
                     count);
  HIPC(hipGetLastError());
  return PSQ_OK;
}

int psq_neighbors_exact_device(const float* d_embedding, uint64_t n_obs, uint64_t n_dim,
                               uint64_t n_neighbors, int32_t metric, uint32_t* d_indptr,
                               uint32_t* d_indices, float* d_dists, uint64_t cap,
                               uint64_t* nnz_out, void* stream) {
  g_err.clear();
  LOCK_DEVICE();
  return neighbors_exact_device_impl(d_embedding, n_obs, n_dim, n_neighbors, metric, d_indptr,
                                     d_indices, d_dists, cap, nnz_out,
                                     static_cast<hipStream_t>(stream));
}

int psq_neighbors_exact(const float* embedding, uint64_t n_obs, uint64_t n_dim,
                        uint64_t n_neighbors, int32_t metric, uint32_t* indptr, uint32_t* indices,
                        float* dists, uint64_t cap, uint64_t* nnz_out, int32_t device) {
  g_err.clear();
  TRY(neighbors_exact_check(n_obs, n_dim, n_neighbors, metric, cap));
  if (!embedding || !indptr || !nnz_out) return fail(PSQ_E_INVALID, "NULL pointer argument");
  const uint64_t kk = std::min<uint64_t>(n_neighbors, n_obs) - 1;
  if (kk == 0) {
    for (uint64_t i = 0; i <= n_obs; ++i) indptr[i] = 0;
    *nnz_out = 0;
    return PSQ_OK;
  }
  if (!indices || !dists) return fail(PSQ_E_INVALID, "NULL pointer argument");
  TRY(set_device(device));
  LOCK_DEVICE();
  Workspace* ws;
  TRY(workspace(&ws));
  float *d_x, *d_dist;
  uint32_t *d_ip, *d_idx;
  const size_t cnt = (size_t)n_obs * kk;
  TRY(wbuf(ws, "ne_data", (size_t)n_obs * n_dim, &d_x));
  TRY(wbuf(ws, "ne_indptr", (size_t)n_obs + 1, &d_ip));
  TRY(wbuf(ws, "ne_indices", cnt, &d_idx));
  TRY(wbuf(ws, "ne_dists", cnt, &d_dist));
  HIPC(hipMemcpy(d_x, embedding, (size_t)n_obs * n_dim * sizeof(float), hipMemcpyHostToDevice));
  uint64_t nnz = 0;
  TRY(neighbors_exact_device_impl(d_x, n_obs, n_dim, n_neighbors, metric, d_ip, d_idx, d_dist, cnt,
                                  &nnz, nullptr));
  HIPC(hipMemcpy(indptr, d_ip, (size_t)(n_obs + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(indices, d_idx, cnt * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(dists, d_dist, cnt * sizeof(float), hipMemcpyDeviceToHost));
  *nnz_out = nnz;
  return PSQ_OK;
}

// ── fuzzy connectivities ──
struct psq_fuzzy_plan {
  int device = 0;
  uint64_t n = 0, nnz = 0;
  uint32_t* ip = nullptr;
  uint32_t* ix = nullptr;
  float* w = nullptr;
};

int psq_fuzzy_connectivities_device(const uint32_t* d_indptr, const uint32_t* d_indices,
                                    const float* d_dists, uint64_t n_obs, uint32_t* d_out_indptr,
                                    uint32_t* d_out_indices, float* d_out_weights, uint64_t cap,
                                    uint64_t* nnz_out, void* stream) {
  g_err.clear();
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!d_indptr || !nnz_out) return fail(PSQ_E_INVALID, "NULL pointer argument");
  if (n_obs >= (1ull << 32)) return fail(PSQ_E_UNSUPPORTED, "n_obs must be < 2^32");
  uint64_t m = 0;
  TRY(fuzzy_input_nnz(d_indptr, n_obs, &m, s));
  if (m > 0 && (!d_indices || !d_dists)) return fail(PSQ_E_INVALID, "NULL pointer argument");
  LOCK_DEVICE();
  Workspace* ws;
  TRY(workspace(&ws));
  FuzzyRun R;
  if (n_obs > 0) TRY(fuzzy_core(ws, d_indptr, d_indices, d_dists, (uint32_t)n_obs, m, s, &R));
  *nnz_out = R.total;
  if (!d_out_indptr || (R.total > 0 && (!d_out_indices || !d_out_weights)))
    return fail(PSQ_E_INVALID, "NULL output pointer");
  if (cap < R.total)
    return fail(PSQ_E_INVALID, "outputs must hold %llu entries (cap %llu)",
                (unsigned long long)R.total, (unsigned long long)cap);
  if (n_obs == 0) {
    TRY(fill_u32(d_out_indptr, 1, 0, s));
  } else {
    HIPC(hipMemcpyAsync(d_out_indptr, R.optr, (n_obs + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    TRY(fuzzy_emit(R, d_out_indices, d_out_weights, s));
  }
  HIPC(hipStreamSynchronize(s));
  return PSQ_OK;
}

int psq_fuzzy_plan_new(const uint32_t* indptr, uint64_t indptr_len, const uint32_t* indices,
                       const float* dists, uint64_t nnz_in, uint64_t n_obs, int32_t device,
                       psq_fuzzy_plan** plan, uint64_t* nnz_out) {
  g_err.clear();
  if (!plan || !nnz_out || !indptr) return fail(PSQ_E_INVALID, "NULL pointer argument");
  *plan = nullptr;
  if (indptr_len != n_obs + 1) return fail(PSQ_E_INVALID, "indptr must have n_obs + 1 entries");
  if (n_obs >= (1ull << 32)) return fail(PSQ_E_UNSUPPORTED, "n_obs must be < 2^32");
  if (indptr[0] != 0) return fail(PSQ_E_INVALID, "indptr[0] must be 0 (got %u)", indptr[0]);
  for (uint64_t i = 0; i < n_obs; ++i)
    if (indptr[i + 1] < indptr[i]) return fail(PSQ_E_INVALID, "indptr must be non-decreasing");
  const uint64_t m = indptr[n_obs];
  if (m > nnz_in)
    return fail(PSQ_E_INVALID, "indptr[n_obs] = %llu exceeds the %llu entries given",
                (unsigned long long)m, (unsigned long long)nnz_in);
  if (m > 0 && (!indices || !dists)) return fail(PSQ_E_INVALID, "NULL pointer argument");
  TRY(set_device(device));
  LOCK_DEVICE();
  auto* P = new psq_fuzzy_plan();
  P->device = device;
  P->n = n_obs;
  auto bail = [&](int rc) {
    psq_fuzzy_plan_free(P);
    return rc;
  };
  hipError_t e = hipMalloc(&P->ip, (n_obs + 1) * sizeof(uint32_t));
  if (e != hipSuccess) return bail(fail(PSQ_E_NOMEM, "hipMalloc failed: %s", hipGetErrorString(e)));
  if (n_obs > 0) {
    Workspace* ws;
    int rc = workspace(&ws);
    if (rc != PSQ_OK) return bail(rc);
    uint32_t *d_ip, *d_ix;
    float* d_dx;
    if ((rc = wbuf(ws, "fz_in_ip", (size_t)n_obs + 1, &d_ip)) != PSQ_OK ||
        (rc = wbuf(ws, "fz_in_ix", std::max<uint64_t>(m, 1), &d_ix)) != PSQ_OK ||
        (rc = wbuf(ws, "fz_in_dx", std::max<uint64_t>(m, 1), &d_dx)) != PSQ_OK)
      return bail(rc);
    if ((e = hipMemcpy(d_ip, indptr, (n_obs + 1) * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (m && (e = hipMemcpy(d_ix, indices, m * 4, hipMemcpyHostToDevice)) != hipSuccess) ||
        (m && (e = hipMemcpy(d_dx, dists, m * 4, hipMemcpyHostToDevice)) != hipSuccess))
      return bail(fail(PSQ_E_DEVICE, "hipMemcpy failed: %s", hipGetErrorString(e)));
    FuzzyRun R;
    if ((rc = fuzzy_core(ws, d_ip, d_ix, d_dx, (uint32_t)n_obs, m, nullptr, &R)) != PSQ_OK)
      return bail(rc);
    P->nnz = R.total;
    if (R.total > 0) {
      if ((e = hipMalloc(&P->ix, R.total * sizeof(uint32_t))) != hipSuccess ||
          (e = hipMalloc(&P->w, R.total * sizeof(float))) != hipSuccess)
        return bail(fail(PSQ_E_NOMEM, "hipMalloc failed: %s", hipGetErrorString(e)));
      if ((rc = fuzzy_emit(R, P->ix, P->w, nullptr)) != PSQ_OK) return bail(rc);
    }
    if ((e = hipMemcpy(P->ip, R.optr, (n_obs + 1) * 4, hipMemcpyDeviceToDevice)) != hipSuccess)
      return bail(fail(PSQ_E_DEVICE, "hipMemcpy failed: %s", hipGetErrorString(e)));
  } else if ((e = hipMemset(P->ip, 0, 4)) != hipSuccess) {
    return bail(fail(PSQ_E_DEVICE, "hipMemset failed: %s", hipGetErrorString(e)));
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess)
    return bail(fail(PSQ_E_DEVICE, "hipDeviceSynchronize failed: %s", hipGetErrorString(e)));
  *plan = P;
  *nnz_out = P->nnz;
  return PSQ_OK;
}

int psq_fuzzy_plan_fill(const psq_fuzzy_plan* plan, uint32_t* indptr, uint64_t indptr_len,
                        uint32_t* indices, float* weights, uint64_t cap) {
  g_err.clear();
  if (!plan || !indptr) return fail(PSQ_E_INVALID, "NULL pointer argument");
  if (indptr_len != plan->n + 1 || cap < plan->nnz)
    return fail(PSQ_E_INVALID, "alloc returned arrays of the wrong size");
  if (plan->nnz > 0 && (!indices || !weights)) return fail(PSQ_E_INVALID, "NULL pointer argument");
  TRY(set_device(plan->device));
  HIPC(hipMemcpy(indptr, plan->ip, (plan->n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (plan->nnz > 0) {
    HIPC(hipMemcpy(indices, plan->ix, plan->nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(weights, plan->w, plan->nnz * sizeof(float), hipMemcpyDeviceToHost));
  }
  return PSQ_OK;
}

void psq_fuzzy_plan_free(psq_fuzzy_plan* plan) {
  if (!plan) return;
  int cur = 0;
  const bool have = hipGetDevice(&cur) == hipSuccess;
  if (hipSetDevice(plan->device) == hipSuccess) {
    if (plan->ip) (void)hipFree(plan->ip);
    if (plan->ix) (void)hipFree(plan->ix);
    if (plan->w) (void)hipFree(plan->w);
  }
  if (have) (void)hipSetDevice(cur);
  delete plan;
}

}  // extern "C"
