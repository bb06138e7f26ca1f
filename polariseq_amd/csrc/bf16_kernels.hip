// Precision form of the local join: BF16 MFMA distance tiles with FP32
// accumulation screen every candidate pair; the pairs that pass the screen are
// re-ranked with the exact f32 distance before anything is proposed.
//
// Same iteration structure as the plain form (crates/polariseq-core/src/
// nndescent.rs:912-969). Per cell i (one wave):
//  * the candidate rows (new list, then old list) and x_i are staged in LDS as
//    f32 with coalesced loads;
//  * MFMA operands are the rows centred on x_i (all candidates are near x_i, so
//    the centred vectors are short) rounded to BF16; one
//    v_mfma_f32_16x16x32_bf16 pair per 16 x 16 tile gives the Gram block of new
//    rows x all candidates, and d~(p,q) = |p~|^2 + |q~|^2 - 2 p~.q~ with the
//    norms taken from the diagonal of each tile's own Gram block (the same
//    BF16 values, the same matrix-core arithmetic);
//  * a pair passes the screen when d~ < slack * thr of either endpoint; passing
//    pairs are listed in LDS and, 64 at a time, get their exact f32 sq_dist
//    (simd.rs:45-62 order) from the staged rows, which is compared with the
//    exact thresholds and proposed as in the plain form.
// Keys in the buckets, the rows and the graph are therefore exact, and the
// plain merge applies. Rows keep 2k entries by default (the graph takes the
// first k), so each row's top 2k are ranked in FP32.
#include "psq_kernels.hpp"

namespace psq {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));


// 8 centred dimensions [t0, t0 + 8) of an LDS row, rounded to BF16, with x_i's
// 8 dimensions in registers (xr, from load8).
// t0 is a multiple of 8 and rows are 16-byte aligned: two 16-byte reads per
// operand, so the 16 rows of a tile, 52 floats apart, fill all 64 banks.
// Reads stop at rs, the floats staged per row (d rounded up to 4); the staged
// floats past d are the rows' zero padding, in the row and in x_i alike, so
// every dimension past d centres to 0 without a test.
__device__ inline bf16x8 centred_bf16x8_r(const float* row, const float (&xr)[8], uint32_t t0,
                                          uint32_t rs) {
  float r[8];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t t = t0 + 4 * h;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < rs) a = *reinterpret_cast<const float4*>(row + t);
    r[4 * h] = a.x; r[4 * h + 1] = a.y; r[4 * h + 2] = a.z; r[4 * h + 3] = a.w;
  }
  bf16x8 v;
#pragma unroll
  for (int u = 0; u < 8; ++u) v[u] = (__bf16)(r[u] - xr[u]);
  return v;
}
__device__ inline void load8(const float* p, uint32_t t0, uint32_t rs, float (&out)[8]) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t t = t0 + 4 * h;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < rs) a = *reinterpret_cast<const float4*>(p + t);
    out[4 * h] = a.x; out[4 * h + 1] = a.y; out[4 * h + 2] = a.z; out[4 * h + 3] = a.w;
  }
}

// One wave (one block) per cell. LDS: pb_key[PB_CAP] u64 | rows[R + 1][RS] f32
// (row R is x_i) | ids[R] | thr[R] | nrm[R] | pb_row[PB_CAP] | sv[320] (screened
// pairs: < 64 held between tiles + 256 per tile; the host sizes it). A block
// holds R candidate rows (a multiple of 16); a cell with more is left to a
// launch with more (jc). d <= 64 (two MFMA steps of depth 32).
template <int D>
__global__ __launch_bounds__(64) void k_join_bf16(
    const float* __restrict__ Xp, uint32_t n, uint32_t d, uint32_t dp, uint32_t RS,
    const uint32_t* __restrict__ newl, const uint32_t* __restrict__ oldl,
    const uint32_t* __restrict__ ncnt, uint32_t m, uint32_t R, JoinCells jc,
    const float* __restrict__ thr, float slack, uint64_t* __restrict__ slots,
    uint8_t* __restrict__ touched, uint32_t S, uint64_t seed, uint64_t iter,
    uint32_t* __restrict__ rec) {
  extern __shared__ __attribute__((aligned(16))) uint64_t lds64[];
  uint64_t* pb_key = lds64;
  float* rows = reinterpret_cast<float*>(lds64 + PB_CAP);
  float* xi = rows + R * RS;
  uint32_t* s_id = reinterpret_cast<uint32_t*>(xi + RS);
  float* s_thr = reinterpret_cast<float*>(s_id + R);
  float* s_nrm = s_thr + R;
  uint32_t* pb_row = reinterpret_cast<uint32_t*>(s_nrm + R);
  uint32_t* sv = pb_row + PB_CAP;
  const uint32_t l = (uint32_t)lane_id();
  const SlotHash sh = slot_hash_of(seed, iter, S);
  uint64_t pairs = 0, exact = 0, sent = 0, rows_in = 0;
  uint32_t i;
  if (!join_cell_of_block(jc, n, i)) return;
  [&] {
    uint32_t nn, no;
    if (!stage_cell<D>(Xp, d, dp, RS, newl, oldl, ncnt, m, thr, i, R, s_id, s_thr, rows,
                       xi, nn, no))
      return;
    const uint32_t nc = nn + no;
    rows_in += nc;
    // floats staged per row: d rounded up to 4 (the padding is zero in every row)
    const uint32_t rsv = D > 0 ? (uint32_t)(D + 3) / 4 * 4 : (d + 3) / 4 * 4;
    __syncthreads();

    uint32_t pcnt = 0, scnt = 0;
    // Exact distances of the listed pairs, 64 at a time (all of them when `all`).
    auto drain = [&](bool all) {
      if (scnt < (all ? 1u : (uint32_t)WAVE)) return;
      __syncthreads();
      uint32_t done = 0;
      while (done < scnt && (all || done + WAVE <= scnt)) {
        const uint32_t e = done + l;
        bool pa = false, pb = false;
        uint32_t ia = 0, ib = 0;
        float dist = 0.0f;
        exact += min(scnt - done, (uint32_t)WAVE);
        if (e < scnt) {
          const uint32_t pr = sv[e], a = pr & 0xFFFF, b = pr >> 16;
          ia = s_id[a];
          ib = s_id[b];
          dist = lds_sq_dist<D>(rows + __umul24(a, RS), rows + __umul24(b, RS), d);
          pa = dist < s_thr[a];
          pb = dist < s_thr[b];
        }
        pb_append(pb_row, pb_key, pcnt, pa, ia, make_key(dist, ib, 1), sent);
        pb_flush(pb_row, pb_key, pcnt, false, slots, touched, S, sh);
        pb_append(pb_row, pb_key, pcnt, pb, ib, make_key(dist, ia, 1), sent);
        pb_flush(pb_row, pb_key, pcnt, false, slots, touched, S, sh);
        done += WAVE;
      }
      if (all) {
        scnt = 0;
        return;
      }
      const uint32_t rest = scnt - done;  // < 64
      const uint32_t keep = l < rest ? sv[done + l] : 0u;
      __syncthreads();
      if (l < rest) sv[l] = keep;
      __syncthreads();
      scnt = rest;
    };

    const uint32_t rtiles = (nn + 15) / 16, ctiles = (nc + 15) / 16;
    const uint32_t lr = l % 16, lk = (l / 16) * 8;
    // this lane's dimensions of x_i for both MFMA steps, read once
    float xr0[8], xr1[8];
    load8(xi, lk, rsv, xr0);
    load8(xi, 32 + lk, rsv, xr1);
    // Column tiles outermost: a tile's operands are built once per cell. In the
    // MFMA layout lane l holds row l % 16, dimensions (l / 16) * 8 + [0, 8) of
    // both operands, so a column tile's operands are also the A operands of the
    // row tile of the same rows: kept for row tile 0, built again for later row
    // tiles. Row tile 0 is never below the diagonal, and a
    // row tile rt > ct is (nn > 16 rt > ct * 16 + 15), so every tile processed
    // has rt <= ct and row tile rt's operands and norms exist when it is used.
    // A row's norm |p~|^2 is the diagonal of its column tile's Gram block: lane l
    // holds outputs (4 (l / 16) + v, l % 16), the diagonal at v = dv.
    const int dv = (int)lr - 4 * (int)(l / 16);
    bf16x8 t0a = {}, t0b = {};  // operands of row tile 0
    for (uint32_t ct = 0; ct < ctiles; ++ct) {
      const float* brow = rows + __umul24(ct * 16 + lr, RS);
      const bf16x8 b0 = centred_bf16x8_r(brow, xr0, lk, rsv);
      const bf16x8 b1 = centred_bf16x8_r(brow, xr1, 32 + lk, rsv);
      {
        f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
        g = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0, b0, g, 0, 0, 0);
        g = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b1, b1, g, 0, 0, 0);
        const float gd = dv == 0 ? g[0] : (dv == 1 ? g[1] : (dv == 2 ? g[2] : g[3]));
        if (dv >= 0 && dv < 4) s_nrm[ct * 16 + lr] = gd;
      }
      if (ct == 0) {
        t0a = b0;
        t0b = b1;
      }
      __syncthreads();  // this tile's norms, written by other lanes
      const uint32_t c = ct * 16 + lr;
      const uint32_t cid = s_id[c];
      const float cnrm = s_nrm[c], cthr = slack * s_thr[c];
      const bool cval = c < nc, cnew = c < nn;
      for (uint32_t rt = 0; rt < rtiles && rt <= ct; ++rt) {
        // skip tiles entirely below the new x new diagonal
        if (ct * 16 + 15 < rt * 16 && ct * 16 + 15 < nn) continue;
        bf16x8 a0 = t0a, a1 = t0b;
        if (rt > 0) {
          const float* arow = rows + __umul24(rt * 16 + lr, RS);
          a0 = centred_bf16x8_r(arow, xr0, lk, rsv);
          a1 = centred_bf16x8_r(arow, xr1, 32 + lk, rsv);
        }
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
        // the 4 rows of this lane's outputs (16-byte aligned, all below R)
        const uint32_t rb = rt * 16 + (l / 16) * 4;
        const uint4 rid = *reinterpret_cast<const uint4*>(s_id + rb);
        const float4 rn = *reinterpret_cast<const float4*>(s_nrm + rb);
        const float4 rth = *reinterpret_cast<const float4*>(s_thr + rb);
        const uint32_t ridv[4] = {rid.x, rid.y, rid.z, rid.w};
        const float rnv[4] = {rn.x, rn.y, rn.z, rn.w}, rthv[4] = {rth.x, rth.y, rth.z, rth.w};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const uint32_t r = rb + v;
          const bool valid = cval && r < nn && (cnew ? c > r : cid != ridv[v]);
          pairs += (uint64_t)__popcll(__ballot(valid));
          // d~ = |p~|^2 + |q~|^2 - 2 p~.q~ against slack x the larger threshold
          const float dt = fmaf(-2.0f, acc[v], rnv[v] + cnrm);
          const bool keep = valid && dt < fmaxf(slack * rthv[v], cthr);
          const uint64_t mask = __ballot(keep);
          if (keep) sv[scnt + (uint32_t)__popcll(mask & ((1ull << l) - 1ull))] = r | (c << 16);
          scnt += (uint32_t)__popcll(mask);
        }
        drain(false);
      }
    }
    drain(true);
    pb_flush(pb_row, pb_key, pcnt, true, slots, touched, S, sh);
  }();
  // pairs, exact, sent and rows_in are wave-uniform (counted from ballots)
  join_record(rec, i, l, (uint32_t)pairs, (uint32_t)rows_in, (uint32_t)exact, (uint32_t)sent);
}

template __global__ void k_join_bf16<0>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                        const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                                        uint32_t, JoinCells, const float*, float, uint64_t*, uint8_t*,
                                        uint32_t, uint64_t, uint64_t, uint32_t*);
template __global__ void k_join_bf16<50>(const float*, uint32_t, uint32_t, uint32_t, uint32_t,
                                         const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                                         uint32_t, JoinCells, const float*, float, uint64_t*, uint8_t*,
                                         uint32_t, uint64_t, uint64_t, uint32_t*);

}  // namespace psq
