// Fused acoustic stage for gfx950 with the WADG projection on FP32 MFMA.
//
// One workgroup (4 waves) owns a tile of TB = 16 consecutive elements:
//   1  Q_in tile -> LDS (buffer loads, 16 B per lane); prefetched into registers: the thread's
//      res and Q_in row of every tile element, c2 in the MFMA output layout, neighbour traces
//   2  (p, Ur, Us, Ut) as one float4 per node; upwind flux phi at face nodes into registers
//      (the p and U fluxes are phi and -phi)
//   3  phi -> LDS [face node][element]; volume terms (4-nnz barycentric operators, 16 float4
//      LDS reads per node) into registers
//   4  volume terms -> rhs (aliasing the (p, U) tile)
//   5  lift: P_f = LIFT_f . phi_f per face on MFMA 16x16x4 (M = nodes, N = 16 elements,
//      K = face nodes); rhs_p += sum_f P_f, rhs_U -= sum_f n_f P_f.  LIFT = EEL . blkdiag(L0)
//      is applied densely: on the matrix cores the factored sparse form saves nothing
//   6  S = (Vq . rhs_p) * c2      MFMA 16x16x4, M = cubature points, N = 16 elements, K = Np
//   7  rhs_p = Pq . S             MFMA 16x16x4, M = nodes, N = 16 elements, K = Nq
//   8  LSRK: res = fa res + fdt rhs (+ source); Q_out = Q_in + fb res (coalesced rows)
// Thread mappings are fixed per thread (node, face node or state row) and loop over the
// tile's elements, so operator rows are loaded once per tile and no index is divided.
// LDS regions are reused across phases (Q_in tile -> face flux -> S; (p, U) -> rhs) to stay
// at 30 KB per workgroup (N = 5).
// With PROJ_U the velocity fields go through Pq . Vq as well (N = 64 columns).
//
// v_mfma_f32_16x16x4f32 lane layout (checked on gfx950): A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k, D[4*(l/16) + r][l%16] in register r of lane l.
#include <algorithm>

#include "bbw_fused_tile.h"

namespace bbw {

// workgroups per CU the register allocation is sized for (LDS allows 5 at N = 5)
#ifndef BBW_FUSED_WG_PER_CU
#define BBW_FUSED_WG_PER_CU 3
#endif

template <int N, bool PROJ_U>
__global__ __launch_bounds__(256, BBW_FUSED_WG_PER_CU) void k_fused_mfma(FusedMesh m, FusedOps o, FusedShot sh,
                                                                         SrcArgs src, float fa, float fb,
                                                                         const float *ftime_dev) {
  using L = FusedLayout<N, PROJ_U>;
  extern __shared__ float lds[];
  const int tid = threadIdx.x;
  const int k0 = tile_index(m.xcd_chunk) * L::TB, nk = min(L::TB, m.K - k0);
  if (ftime_dev) sh.ft = *ftime_dev;
  const int fid3 = face_node<N, PROJ_U>(o, tid);
  // vmapP first: the trace gather that depends on it is issued behind the tile loads
  int vm[L::EPT3];
  vm_load<N, PROJ_U>(vm, m, k0, nk, tid);
  TileStage<N, PROJ_U> st;
  stage_load<N, PROJ_U>(st, m, sh.Qin, k0, nk, tid);
  TilePre<N, PROJ_U> pf;
  tile_prefetch<N, PROJ_U>(pf, m, o, sh, fa, k0, nk, vm, tid);
  stage_store<N, PROJ_U>(st, lds + L::OQ, lds + L::OG, tid);
  __syncthreads();
  tile_shot<N, PROJ_U>(m, o, sh, src, fa, fb, k0, nk, lds, lds + L::OQ, lds + L::OG, vm, pf, fid3, tid,
                       NoHook{}, NoHook{}, NoHook{});
}

// Plain WADG update on MFMA (the three-launch path's rk_update_WADG,
// cpp/okl/WaveKernelsWADG.okl:1760-1887): per 16-element tile, all four fields projected,
// rhs_f <- Pq (Vq rhs_f x c2/J (p) or 1/J (u, v, w)), then the LSRK stage with the source.
// N = 64 columns (4 fields x 16 elements); the same lane-ordered operand blocks as the fused
// stage, fetched one phase ahead.
template <int N>
struct UpdLayout {
  static constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq, TB = 16, NpNf = 4 * Np;
  static constexpr int NQT = (Nq + 15) / 16, NQP = 16 * NQT, NPT = (Np + 15) / 16;
  static constexpr int KTV = (Np + 3) / 4, KTP = NQP / 4, QTW = (NQT + 3) / 4;
  static constexpr int KBV = (KTV + 3) / 4, KBP = (KTP + 3) / 4;
  static constexpr int RS0 = 4 * Np > 3 * Np + 4 * KTV ? 4 * Np : 3 * Np + 4 * KTV;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  static constexpr bool PF = NpNf <= 256;
  // rhs rows [TB][RS] | S [4 fields][NQP][16] | source slots [TB] | one scratch word
  static constexpr size_t floats = (size_t)TB * RS + 4 * NQP * 16 + TB + 1;
};

// A tile's rows [TB][NW] (row stride ROW bytes in HBM) held in registers between their loads
// and their LDS stores (rows of stride RS), so the kernel's other loads can be issued in
// between and every store is unconditional.  Lanes past a row read through an out-of-range
// offset (0, no memory traffic) and store to a scratch word: a store under a lane branch lets
// the compiler sink its load into the branch and wait for each round in turn.
template <int NW, int TB>
struct TileRows {
  static constexpr int RR = (NW + 255) / 256;
  float h[RR][TB];
  __device__ __forceinline__ void load(brsrc r, unsigned ROW, int tid) {
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      const int rr = tid + 256 * k;
#pragma unroll
      for (int e = 0; e < TB; ++e) h[k][e] = bld(r, rr < NW ? (unsigned)rr * 4u + e * ROW : 0x80000000u);
    }
  }
  template <int RS>
  __device__ __forceinline__ void store(float *sR, float *sScr, int tid) const {
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      const int rr = tid + 256 * k;
      float *dst = rr < NW ? sR + rr : sScr;
      const int step = rr < NW ? RS : 0;
#pragma unroll
      for (int e = 0; e < TB; ++e) dst[e * step] = h[k][e];
    }
  }
};

// LSRK stage over a tile whose rhs rows [TB][NW] sit in LDS (stride RS): res = fa res +
// fdt (rhs + source), Q += fb res, for rows of NW words (fields of Np nodes) in HBM with row
// stride ROW bytes.  Every res and Q load of the thread is issued before the first store; lanes
// past a row use an out-of-range offset (the loads read 0, the stores are dropped).  wfield(f)
// is the source weight of field f.
template <int NW, int RS, int TB, int Np, class WF>
__device__ __forceinline__ void lsrk_rows(brsrc rRs, brsrc rQ, unsigned ROW, const float *sR,
                                          const int *sSlot, bool has_src, const SrcArgs &src,
                                          float ft, float fa, float fb, float fdt, int tid, WF &&wfield) {
  constexpr int RR = (NW + 255) / 256;
  float rv[RR][TB], qv[RR][TB];
#pragma unroll
  for (int k = 0; k < RR; ++k) {
    const int rr = tid + 256 * k;
#pragma unroll
    for (int e = 0; e < TB; ++e) {
      const unsigned g = rr < NW ? (unsigned)rr * 4u + e * ROW : 0x80000000u;
      rv[k][e] = bld(rRs, g);
      qv[k][e] = bld(rQ, g);
    }
  }
#pragma unroll
  for (int k = 0; k < RR; ++k) {
    const int rr = tid + 256 * k;
    const int rc = rr < NW ? rr : 0;   // LDS reads of lanes past the row stay in the tile
    const int fr = rc / Np, nr = rc - fr * Np;
    const float wsel = wfield(fr);
#pragma unroll
    for (int e = 0; e < TB; ++e) {
      float v = sR[e * RS + rc];
      if (has_src) {
        const int slot = sSlot[e];
        if (slot >= 0) v += wsel * ft * src.coef[(size_t)slot * Np + nr];
      }
      const unsigned g = rr < NW ? (unsigned)rr * 4u + e * ROW : 0x80000000u;
      const float x = fa * rv[k][e] + fdt * v;
      bst(rRs, x, g);
      bst(rQ, qv[k][e] + fb * x, g);
    }
  }
}

template <int N>
__global__ __launch_bounds__(256, N <= 4 ? 4 : N == 5 ? 3 : 2) void k_update_wadg_mfma(int K, FusedOps o, const float *__restrict__ c2q,
                                                             const float *__restrict__ Jq, float fa,
                                                             float fb, float fdt, SrcArgs src,
                                                             const float *ftime_dev, float ftime,
                                                             const float *__restrict__ rhs,
                                                             float *__restrict__ res,
                                                             float *__restrict__ Q) {
  using L = UpdLayout<N>;
  constexpr int Np = L::Np, Nq = L::Nq, TB = L::TB, NpNf = L::NpNf, RS = L::RS, NQP = L::NQP;
  constexpr int KTV = L::KTV, KTP = L::KTP, KBV = L::KBV, KBP = L::KBP;
  constexpr unsigned ROW = NpNf * 4u;
  extern __shared__ float lds[];
  float *sR = lds;                         // [TB][RS]
  float *S = sR + TB * RS;                 // [4][NQP][16]
  int *sSlot = (int *)(S + 4 * NQP * 16);  // [TB]
  float *sScr = (float *)(sSlot + TB);     // written by lanes past a row, never read
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * NpNf;
  const unsigned tbytes = (unsigned)nk * ROW, qbytes = (unsigned)nk * Nq * 4u;
  const brsrc rRh = mkrsrc(rhs + base, tbytes), rRs = mkrsrc(res + base, tbytes);
  const brsrc rQ = mkrsrc(Q + base, tbytes), rC = mkrsrc(c2q + (size_t)k0 * Nq, qbytes);
  const brsrc rJ = mkrsrc(Jq ? Jq + (size_t)k0 * Nq : nullptr, Jq ? qbytes : 0u);
  const brsrc rAV = mkrsrc(o.VqBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rAP = mkrsrc(o.PqBlk, (unsigned)L::NPT * KBP * 1024u);
  static_assert(TB == 16, "the per-tile source table is built for 16-element tiles");
  const int2 stile = src_tile(src, k0, nk);
  const bool has_src = stile.y != 0;

  // ---- 1: rhs tile -> LDS; res and Q rows, c2 and 1/J (MFMA output layout), Vq blocks.
  // With fa = 0 (the first LSRK45 stage) res is not read: fa res = 0 for any finite res.
  // The rhs rows are stored to LDS after every load of the phase is issued.
  const bool read_res = fa != 0.f;
  float rreg[L::PF ? TB : 1], qreg[L::PF ? TB : 1];
  TileRows<NpNf, TB> rows;
  if constexpr (L::PF) {
#pragma unroll
    for (int e = 0; e < TB; ++e) {
      rows.h[0][e] = bld(rRh, tid * 4u + e * ROW);
      qreg[e] = bld(rQ, tid * 4u + e * ROW);
    }
    if (read_res) {
#pragma unroll
      for (int e = 0; e < TB; ++e) rreg[e] = bld(rRs, tid * 4u + e * ROW);
    } else {
#pragma unroll
      for (int e = 0; e < TB; ++e) rreg[e] = 0.f;
    }
  } else {
    rows.load(rRh, ROW, tid);
  }
  constexpr int KPAD = L::RS0 - NpNf;
  static_assert(TB * KPAD <= 256, "one padding word per thread");
  if (KPAD > 0 && tid < TB * KPAD) sR[(tid / KPAD) * RS + NpNf + tid % KPAD] = 0.f;
  if (has_src && tid < TB) sSlot[tid] = src_slot_tile(src, stile, k0, tid);
  float c2r[L::QTW][4], jr[L::QTW][4];
  f4 aV[L::QTW][KBV];
  const unsigned co = (unsigned)(col * Nq + 4 * kr) * 4u;
  // unconditional (a wave past the last tile loads the last one and does not use it): loads
  // under a wave test leave the compiler unsure how many are in flight where the paths join,
  // and its waits then cover later prefetches as well
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t < L::NQT ? wave + 4 * t : L::NQT - 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      c2r[t][r] = bld(rC, co + (qt * 16 + r) * 4u);
      jr[t][r] = Jq ? 1.f / bld(rJ, co + (qt * 16 + r) * 4u) : 1.f;
    }
#pragma unroll
    for (int kb = 0; kb < KBV; ++kb) aV[t][kb] = bld_blk<KTV>(rAV, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
  }
  rows.template store<RS>(sR, sScr, tid);
  __syncthreads();

  // ---- 2: S_f = (Vq . rhs_f) x c2/J (p) or 1/J (M = cubature points, N = elements, K = nodes)
  f4 aP[KBP];
  {
    const int mt0 = wave < L::NPT ? wave : L::NPT - 1;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<KTP>(rAP, lane * 16u, (unsigned)(mt0 * KBP + kb) * 1024u, kb);
  }
  {
    const float *Bp = sR + col * RS + kr;
#pragma unroll
    for (int t = 0; t < L::QTW; ++t) {
      const int qt = wave + 4 * t;
      if (qt >= L::NQT) break;
      f4 acc[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < KTV) {
#pragma unroll
            for (int b = 0; b < 4; ++b)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aV[t][kb][sb], Bp[b * Np + kt * 4], acc[b], 0, 0, 0);
          }
        }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        const bool in = q < Nq;
        S[q * 16 + col] = in ? acc[0][r] * c2r[t][r] * jr[t][r] : 0.f;
#pragma unroll
        for (int b = 1; b < 4; ++b) S[(b * NQP + q) * 16 + col] = in ? acc[b][r] * jr[t][r] : 0.f;
      }
    }
  }
  __syncthreads();

  // ---- 3: rhs_f = Pq . S_f (M = nodes, K = cubature points)
  for (int mt = wave; mt < L::NPT; mt += 4) {
    if (mt != wave) {
#pragma unroll
      for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<KTP>(rAP, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
    }
    f4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
    const float *Bp = S + kr * 16 + col;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb)
#pragma unroll
      for (int sb = 0; sb < 4; ++sb) {
        const int kt = 4 * kb + sb;
        if (kt < KTP) {
#pragma unroll
          for (int b = 0; b < 4; ++b)
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[kb][sb], Bp[(b * NQP + kt * 4) * 16], acc[b], 0, 0, 0);
        }
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = mt * 16 + 4 * kr + r;
      if (n < Np) {
#pragma unroll
        for (int b = 0; b < 4; ++b) sR[col * RS + b * Np + n] = acc[b][r];
      }
    }
  }
  __syncthreads();

  // ---- 4: LSRK stage: res = fa res + fdt (rhs + source); Q += fb res
  const float ft = ftime_dev ? *ftime_dev : ftime;
  if constexpr (L::PF) {
    if (tid < NpNf) {
      const int fr = tid / Np, nr = tid - fr * Np;
      const float wsel = w4(src, fr);
#pragma unroll
      for (int e = 0; e < TB; ++e) {
        float v = sR[e * RS + tid];
        if (has_src) {
          const int slot = sSlot[e];
          if (slot >= 0) v += wsel * ft * src.coef[(size_t)slot * Np + nr];
        }
        const float x = fa * rreg[e] + fdt * v;
        bst(rRs, x, tid * 4u + e * ROW);
        bst(rQ, qreg[e] + fb * x, tid * 4u + e * ROW);
      }
    }
  } else {
    lsrk_rows<NpNf, RS, TB, Np>(rRs, rQ, ROW, sR, sSlot, has_src, src, ft, fa, fb, fdt, tid,
                                [&](int f) { return w4(src, f); });
  }
}

template <int N>
static hipError_t update_wadg_mfma_n(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                     float ft, const SrcArgs &s) {
  using L = UpdLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_update_wadg_mfma<N>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_update_wadg_mfma<N>, dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream,
                     c->K, fused_ops(c), c->c2q, c->Jq, fa, fb, fdt, s, ftd, ft, c->rhs, c->res,
                     c->Q[c->cur]);
  return hipGetLastError();
}

// The same update for orders 7..9, whose cubature (214, 729, 1000 points) does not fit LDS as a
// whole: the points are taken in chunks of QC; per chunk S_f = (Vq_chunk rhs_f) x c2/J or 1/J
// goes through LDS and rhs_f accumulates Pq_chunk S_f in registers over the chunks.
template <int N>
struct UpdChunkLayout {
  static constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq, TB = 16, NpNf = 4 * Np;
  static constexpr int NQT = (Nq + 15) / 16, NPT = (Np + 15) / 16, MTW = (NPT + 3) / 4;
  static constexpr int KTV = (Np + 3) / 4, KBV = (KTV + 3) / 4;
  static constexpr int QCT = 4, QC = 16 * QCT;           // q tiles and points per chunk
  static constexpr int NCH = (NQT + QCT - 1) / QCT;
  static constexpr int RS0 = 4 * Np > 3 * Np + 4 * KTV ? 4 * Np : 3 * Np + 4 * KTV;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  // rhs rows [TB][RS] | S [4 fields][QC][16] | source slots [TB] | one scratch word
  static constexpr size_t floats = (size_t)TB * RS + 4 * QC * 16 + TB + 1;
};

template <int N>
__global__ __launch_bounds__(256) void k_update_wadg_chunk(int K, const float *__restrict__ VqBlk,
                                                           const float *__restrict__ PqBlk,
                                                           const float *__restrict__ c2q,
                                                           const float *__restrict__ Jq, float fa,
                                                           float fb, float fdt, SrcArgs src,
                                                           const float *ftime_dev, float ftime,
                                                           const float *__restrict__ rhs,
                                                           float *__restrict__ res,
                                                           float *__restrict__ Q) {
  using L = UpdChunkLayout<N>;
  constexpr int Np = L::Np, Nq = L::Nq, TB = L::TB, NpNf = L::NpNf, RS = L::RS, QC = L::QC;
  constexpr unsigned ROW = NpNf * 4u;
  extern __shared__ float lds[];
  float *sR = lds;                         // [TB][RS]
  float *S = sR + TB * RS;                 // [4][QC][16]
  int *sSlot = (int *)(S + 4 * QC * 16);   // [TB]
  float *sScr = (float *)(sSlot + TB);
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * NpNf;
  const unsigned tbytes = (unsigned)nk * ROW, qbytes = (unsigned)nk * Nq * 4u;
  const brsrc rRh = mkrsrc(rhs + base, tbytes), rRs = mkrsrc(res + base, tbytes);
  const brsrc rQ = mkrsrc(Q + base, tbytes), rC = mkrsrc(c2q + (size_t)k0 * Nq, qbytes);
  const brsrc rJ = mkrsrc(Jq ? Jq + (size_t)k0 * Nq : nullptr, Jq ? qbytes : 0u);
  const brsrc rAV = mkrsrc(VqBlk, (unsigned)L::NQT * L::KBV * 1024u);
  const brsrc rAP = mkrsrc(PqBlk, (unsigned)L::NPT * L::NQT * 1024u);
  static_assert(TB == 16, "the per-tile source table is built for 16-element tiles");
  const int2 stile = src_tile(src, k0, nk);
  const bool has_src = stile.y != 0;

  // ---- 1: rhs tile -> LDS (zero padding past the last field for the K tail)
  {
    TileRows<NpNf, TB> rows;
    rows.load(rRh, ROW, tid);
    constexpr int KPAD = L::RS0 - NpNf;
    static_assert(TB * KPAD <= 256, "one padding word per thread");
    if (KPAD > 0 && tid < TB * KPAD) sR[(tid / KPAD) * RS + NpNf + tid % KPAD] = 0.f;
    if (has_src && tid < TB) sSlot[tid] = src_slot_tile(src, stile, k0, tid);
    rows.template store<RS>(sR, sScr, tid);
  }
  __syncthreads();

  f4 out[L::MTW][4];
#pragma unroll
  for (int i = 0; i < L::MTW; ++i)
#pragma unroll
    for (int b = 0; b < 4; ++b) out[i][b] = f4{0.f, 0.f, 0.f, 0.f};
  const unsigned co = (unsigned)(col * Nq + 4 * kr) * 4u;
  for (int ch = 0; ch < L::NCH; ++ch) {
    // ---- 2: S_f for the chunk's q tiles, one per wave
    {
      const int qt = ch * L::QCT + wave;       // QCT = 4 tiles per chunk, one per wave
      f4 acc[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
      if (qt < L::NQT) {
        const float *Bp = sR + col * RS + kr;
        for (int kb = 0; kb < L::KBV; ++kb) {
          const f4 a = bld4_s(rAV, lane * 16u, (unsigned)(qt * L::KBV + kb) * 1024u);
#pragma unroll
          for (int sb = 0; sb < 4; ++sb) {
            const int kt = 4 * kb + sb;
            if (kt < L::KTV) {
#pragma unroll
              for (int b = 0; b < 4; ++b)
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[sb], Bp[b * Np + kt * 4], acc[b], 0, 0, 0);
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        const bool in = qt < L::NQT && q < Nq;
        const unsigned qo = co + (unsigned)(qt * 16 + r) * 4u;
        const float c2 = in ? bld(rC, qo) : 0.f;
        const float jr = in ? (Jq ? 1.f / bld(rJ, qo) : 1.f) : 0.f;
        const int ql = wave * 16 + 4 * kr + r;
        S[ql * 16 + col] = acc[0][r] * c2 * jr;
#pragma unroll
        for (int b = 1; b < 4; ++b) S[(b * QC + ql) * 16 + col] = acc[b][r] * jr;
      }
    }
    __syncthreads();
    // ---- 3: rhs_f += Pq_chunk S_f (M = nodes, K = the chunk's points)
#pragma unroll
    for (int i = 0; i < L::MTW; ++i) {
      const int mt = wave + 4 * i;
      if (mt < L::NPT) {
        const float *Bp = S + kr * 16 + col;
#pragma unroll
        for (int t = 0; t < L::QCT; ++t) {
          const int qt = ch * L::QCT + t;
          if (qt < L::NQT) {
            const f4 a = bld4_s(rAP, lane * 16u, (unsigned)(mt * L::NQT + qt) * 1024u);
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) {
              const int k = t * 16 + sb * 4;
#pragma unroll
              for (int b = 0; b < 4; ++b)
                out[i][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[sb], Bp[(b * QC + k) * 16], out[i][b], 0, 0, 0);
            }
          }
        }
      }
    }
    __syncthreads();
  }
  // ---- 4: rhs rows in LDS, then the LSRK stage
#pragma unroll
  for (int i = 0; i < L::MTW; ++i) {
    const int mt = wave + 4 * i;
    if (mt < L::NPT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int b = 0; b < 4; ++b) sR[col * RS + b * Np + n] = out[i][b][r];
        }
      }
    }
  }
  __syncthreads();
  const float ft = ftime_dev ? *ftime_dev : ftime;
  lsrk_rows<NpNf, RS, TB, Np>(rRs, rQ, ROW, sR, sSlot, has_src, src, ft, fa, fb, fdt, tid,
                              [&](int f) { return w4(src, f); });
}

template <int N>
static hipError_t update_wadg_chunk_n(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                      float ft, const SrcArgs &s) {
  using L = UpdChunkLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_update_wadg_chunk<N>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_update_wadg_chunk<N>, dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh,
                     c->stream, c->K, c->VqBlk, c->PqBlk, c->c2q, c->Jq, fa, fb, fdt, s, ftd, ft,
                     c->rhs, c->res, c->Q[c->cur]);
  return hipGetLastError();
}

hipError_t launch_update_wadg_mfma(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                   float ft) {
  SrcArgs s = ctx_src(c);
  switch (c->N) {
    case 1: return update_wadg_mfma_n<1>(c, fa, fb, fdt, ftd, ft, s);
    case 2: return update_wadg_mfma_n<2>(c, fa, fb, fdt, ftd, ft, s);
    case 3: return update_wadg_mfma_n<3>(c, fa, fb, fdt, ftd, ft, s);
    case 4: return update_wadg_mfma_n<4>(c, fa, fb, fdt, ftd, ft, s);
    case 5: return update_wadg_mfma_n<5>(c, fa, fb, fdt, ftd, ft, s);
    case 6: return update_wadg_mfma_n<6>(c, fa, fb, fdt, ftd, ft, s);
    case 7: return update_wadg_chunk_n<7>(c, fa, fb, fdt, ftd, ft, s);
    case 8: return update_wadg_chunk_n<8>(c, fa, fb, fdt, ftd, ft, s);
    case 9: return update_wadg_chunk_n<9>(c, fa, fb, fdt, ftd, ft, s);
  }
  return hipErrorInvalidValue;
}

// Plain volume and surface kernels on 16-element tiles (the three-launch path's
// rk_volume_bern, cpp/okl/WaveKernels.okl:344-474, and rk_surface_bern, :479-647): the same
// buffer addressing and per-thread operand addresses as the fused stage.  The surface kernel
// keeps the reference's lift: L0 per face, then EEL, accumulated into rhs.
template <int N>
struct PlainLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, TB = 16;
  static constexpr int NpNf = 4 * Np, QT = TB * NpNf;
  static constexpr unsigned QB = QT * 4u, QIT = (QB + 4095) / 4096;
  static constexpr int QTP = (int)QIT * 1024;  // the Q tile padded to whole 4 KB rounds
  static constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS, EPTV = (TB + NG - 1) / NG;
  static constexpr int NG3 = 256 / NfpNf, EPT3 = (TB + NG3 - 1) / NG3;
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  // volume: Q tile [TB][4Np] (QTP words) | (p, Ur, Us, Ut) [TB][Np] float4 | vgeo [TB][9] in a
  // 256-word slot
  static constexpr size_t vfloats = (size_t)QTP + QT + 256;
  // surface: Q tile [TB][4Np] (QTP words) then lifted face data float4 [TB][NfpNf] | phi
  // [TB][NfpNf] | fgeo
  static constexpr size_t sfloats = (size_t)cmax(QTP, 4 * TB * NfpNf) + TB * NfpNf + TB * 16;
};

// The Q tile in registers (QIT 4 KB rounds of float4 per workgroup); stored to LDS after the
// kernel's other loads are issued, with unconditional stores (the LDS tile holds whole rounds;
// words past the tile hold zeros through the range check and are never read).
template <int N>
struct PlainQ {
  float4 v[PlainLayout<N>::QIT];
};

template <int N>
__device__ __forceinline__ void plain_q_load(PlainQ<N> &t, const float *Q, int k0, int nk, int tid) {
  using L = PlainLayout<N>;
  const brsrc rQ = mkrsrc(Q + (size_t)k0 * L::NpNf, (unsigned)nk * L::NpNf * 4u);
#pragma unroll
  for (unsigned it = 0; it < L::QIT; ++it) t.v[it] = bld4(rQ, tid * 16u + it * 4096u);
}

template <int N>
__device__ __forceinline__ void plain_q_store(const PlainQ<N> &t, float *sQ, int tid) {
  float4 *sQ4 = (float4 *)sQ;
#pragma unroll
  for (unsigned it = 0; it < PlainLayout<N>::QIT; ++it) sQ4[tid + it * 256u] = t.v[it];
}

template <int N>
__global__ __launch_bounds__(256) void k_volume_t(int K, const float *__restrict__ vgeo,
                                                  const int4 *__restrict__ Dids,
                                                  const float4 *__restrict__ Dvals,
                                                  const float *__restrict__ Q, float *__restrict__ rhs) {
  using L = PlainLayout<N>;
  constexpr int Np = L::Np, NpNf = L::NpNf, TB = L::TB;
  extern __shared__ float lds[];
  float *sQ = lds;
  float4 *sV = (float4 *)(lds + L::QTP);
  float *sG = lds + L::QTP + L::QT;
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const int gv = tid / L::GS, nv = tid - gv * L::GS;
  const bool onv = gv < L::NG && nv < Np;
  // the node's derivative rows are loaded with the tile, not after the barriers (lanes past
  // the nodes read row 0 and return before using it)
  const int nr = nv < Np ? nv : 0;
  const float4 dv = Dvals[nr];
  const int4 D1 = Dids[nr], D2 = Dids[Np + nr], D3 = Dids[2 * Np + nr], D4 = Dids[3 * Np + nr];
  {
    PlainQ<N> qt;
    plain_q_load<N>(qt, Q, k0, nk, tid);
    const float gval = bld(mkrsrc(vgeo + (size_t)k0 * 9, (unsigned)nk * 36u), tid * 4u);
    plain_q_store<N>(qt, sQ, tid);
    sG[tid] = gval;
  }
  __syncthreads();
  if (onv) {
#pragma unroll
    for (int i = 0; i < L::EPTV; ++i) {
      const int e = gv + L::NG * i;
      const float *q = sQ + e * NpNf;
      const float *G = sG + 9 * e;
      const float uu = q[Np + nv], vv = q[2 * Np + nv], ww = q[3 * Np + nv];
      sV[e * Np + nv] = make_float4(q[nv], uu * G[0] + vv * G[1] + ww * G[2],
                                    uu * G[3] + vv * G[4] + ww * G[5], uu * G[6] + vv * G[7] + ww * G[8]);
    }
  }
  __syncthreads();
  if (!onv) return;
  const brsrc rO = mkrsrc(rhs + (size_t)k0 * NpNf, (unsigned)nk * NpNf * 4u);
  const float4 *Vb = sV + gv * Np;
  const float4 *pa1 = Vb + D1.x, *pa2 = Vb + D2.x, *pa3 = Vb + D3.x, *pa4 = Vb + D4.x;
  const float4 *pb1 = Vb + D1.y, *pb2 = Vb + D2.y, *pb3 = Vb + D3.y, *pb4 = Vb + D4.y;
  const float4 *pc1 = Vb + D1.z, *pc2 = Vb + D2.z, *pc3 = Vb + D3.z, *pc4 = Vb + D4.z;
  const float4 *pd1 = Vb + D1.w, *pd2 = Vb + D2.w, *pd3 = Vb + D3.w, *pd4 = Vb + D4.w;
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i) {
    const int e = gv + L::NG * i;
    const int o4 = i * L::NG * Np;
    const float4 a1 = pa1[o4], a2 = pa2[o4], a3 = pa3[o4], a4 = pa4[o4];
    const float4 b1 = pb1[o4], b2 = pb2[o4], b3 = pb3[o4], b4 = pb4[o4];
    const float4 c1 = pc1[o4], c2 = pc2[o4], c3 = pc3[o4], c4 = pc4[o4];
    const float4 d1 = pd1[o4], d2 = pd2[o4], d3 = pd3[o4], d4 = pd4[o4];
    const float p1 = dv.x * a1.x + dv.y * a2.x + dv.z * a3.x + dv.w * a4.x;
    const float p2 = dv.x * b1.x + dv.y * b2.x + dv.z * b3.x + dv.w * b4.x;
    const float p3 = dv.x * c1.x + dv.y * c2.x + dv.z * c3.x + dv.w * c4.x;
    const float p4 = dv.x * d1.x + dv.y * d2.x + dv.z * d3.x + dv.w * d4.x;
    float dU1 = dv.x * (b1.y + c1.z + d1.w);
    dU1 += dv.y * (b2.y + c2.z + d2.w);
    dU1 += dv.z * (b3.y + c3.z + d3.w);
    dU1 += dv.w * (b4.y + c4.z + d4.w);
    float dU2 = dv.x * (a1.y + a1.z + a1.w);
    dU2 += dv.y * (a2.y + a2.z + a2.w);
    dU2 += dv.z * (a3.y + a3.z + a3.w);
    dU2 += dv.w * (a4.y + a4.z + a4.w);
    const float dpdr = .5f * (p2 - p1), dpds = .5f * (p3 - p1), dpdt = .5f * (p4 - p1);
    const float *G = sG + 9 * e;
    const unsigned o = (unsigned)(e * NpNf + nv) * 4u;
    bst(rO, -.5f * (dU1 - dU2), o);
    bst(rO, -(G[0] * dpdr + G[3] * dpds + G[6] * dpdt), o + Np * 4u);
    bst(rO, -(G[1] * dpdr + G[4] * dpds + G[7] * dpdt), o + 2 * Np * 4u);
    bst(rO, -(G[2] * dpdr + G[5] * dpds + G[8] * dpdt), o + 3 * Np * 4u);
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_surface_t(int K, const float *__restrict__ fgeo,
                                                   const int *__restrict__ Fmask,
                                                   const int *__restrict__ vmapP,
                                                   const int *__restrict__ EEL_ids4,
                                                   const float *__restrict__ EEL_vals,
                                                   const int *__restrict__ L0_ids,
                                                   const float *__restrict__ L0_vals,
                                                   const float *__restrict__ Q, float *__restrict__ rhs) {
  using L = PlainLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, NpNf = L::NpNf, TB = L::TB;
  constexpr int EEL_nnz = Dims<N>::EEL_nnz, L0_nnz = Dims<N>::L0_nnz;
  extern __shared__ float lds[];
  float *sQ = lds;                                        // [TB][4Np]       (phases 1-2)
  float4 *sL = (float4 *)lds;                             // [TB][NfpNf]     (phases 3-4)
  float *sPhi = lds + L::cmax(L::QTP, 4 * TB * NfpNf);    // [TB][NfpNf]
  float *sF = sPhi + TB * NfpNf;                          // [TB][16]
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const brsrc rO = mkrsrc(rhs + (size_t)k0 * NpNf, (unsigned)nk * NpNf * 4u);
  const brsrc rT = mkrsrc(Q, (unsigned)K * NpNf * 4u);
  // ---- 1: Q tile, fgeo; the thread's vmapP entries, neighbour traces and rhs rows (all loads
  // issued before the LDS stores)
  PlainQ<N> qt;
  plain_q_load<N>(qt, Q, k0, nk, tid);
  const float fval = bld(mkrsrc(fgeo + (size_t)k0 * 16, (unsigned)nk * 64u), tid * 4u);
  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf;
  const bool on3 = g3 < L::NG3;
  const int f3 = m3 / Nfp;
  const int fid3 = on3 ? Fmask[m3] : 0;
  const FusedMesh mm{K, nullptr, nullptr, vmapP};
  int vm[L::EPT3];
  vm_load<N, false>(vm, mm, k0, nk, tid);
  float tr[L::EPT3][4];
#pragma unroll
  for (int i = 0; i < L::EPT3; ++i)
#pragma unroll
    for (int f = 0; f < 4; ++f) tr[i][f] = bld(rT, (unsigned)vm[i] * 4u + f * Np * 4u);
  const int gv = tid / L::GS, nv = tid - gv * L::GS;
  const bool onv = gv < L::NG && nv < Np;
  float ro[L::EPTV][4];
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i)
#pragma unroll
    for (int f = 0; f < 4; ++f) ro[i][f] = bld(rO, (unsigned)((gv + L::NG * i) * NpNf + f * Np + nv) * 4u);
  plain_q_store<N>(qt, sQ, tid);
  sF[tid] = fval;
  __syncthreads();
  // ---- 2: upwind flux phi = (pjump - Unjump) Fscale / 2 (the U flux is -phi)
  if (on3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) {
        const float *q = sQ + e * NpNf;
        const float *F = sF + 16 * e + 4 * f3;
        const float pM = q[fid3], uM = q[Np + fid3], vM = q[2 * Np + fid3], wM = q[3 * Np + fid3];
        float pjump, Unjump;
        if (vm[i] == fid3 + (k0 + e) * NpNf) {
          pjump = -2.f * pM;
          Unjump = 0.f;
        } else {
          pjump = tr[i][0] - pM;
          Unjump = (tr[i][1] - uM) * F[1] + (tr[i][2] - vM) * F[2] + (tr[i][3] - wM) * F[3];
        }
        sPhi[e * NfpNf + m3] = .5f * (pjump - Unjump) * F[0];
      }
    }
  }
  __syncthreads();
  // ---- 3: L0 per face; stored as (p, U nx, U ny, U nz) for the EEL rows
  if (on3) {
    const int nt = m3 - f3 * Nfp;
    float Lv[L0_nnz];
    const float *Xp[L0_nnz];
    const float *Xb = sPhi + g3 * NfpNf + f3 * Nfp;
#pragma unroll
    for (int j = 0; j < L0_nnz; ++j) {
      Lv[j] = L0_vals[nt + j * Nfp];
      Xp[j] = Xb + L0_ids[nt + j * Nfp];
    }
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) {
        float v = 0.f;
#pragma unroll
        for (int j = 0; j < L0_nnz; ++j) v += Lv[j] * Xp[j][i * L::NG3 * NfpNf];
        const float *F = sF + 16 * e + 4 * f3;
        sL[e * NfpNf + m3] = make_float4(v, -v * F[1], -v * F[2], -v * F[3]);
      }
    }
  }
  __syncthreads();
  // ---- 4: EEL rows, accumulated into rhs
  if (!onv) return;
  float Ev[EEL_nnz];
  const float4 *Ep[EEL_nnz];
  const float *Lb = (const float *)(sL + gv * NfpNf);
#pragma unroll
  for (int j = 0; j < EEL_nnz; ++j) {
    Ev[j] = EEL_vals[nv + j * Np];
    Ep[j] = (const float4 *)(Lb + EEL_ids4[nv + j * Np]);
  }
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i) {
    const int e = gv + L::NG * i;
    float v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f;
#pragma unroll
    for (int j = 0; j < EEL_nnz; ++j) {
      const float4 x = Ep[j][i * L::NG * NfpNf];
      v1 += Ev[j] * x.x;
      v2 += Ev[j] * x.y;
      v3 += Ev[j] * x.z;
      v4 += Ev[j] * x.w;
    }
    const unsigned o = (unsigned)(e * NpNf + nv) * 4u;
    bst(rO, ro[i][0] + v1, o);
    bst(rO, ro[i][1] + v2, o + Np * 4u);
    bst(rO, ro[i][2] + v3, o + 2 * Np * 4u);
    bst(rO, ro[i][3] + v4, o + 3 * Np * 4u);
  }
}

template <int N>
static hipError_t volume_t_n(bbw_ctx *c) {
  using L = PlainLayout<N>;
  const size_t sh = sizeof(float) * L::vfloats;
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_volume_t<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_volume_t<N>, dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream, c->K,
                     c->vgeo, c->Dids, c->Dvals, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

template <int N>
static hipError_t surface_t_n(bbw_ctx *c) {
  using L = PlainLayout<N>;
  const size_t sh = sizeof(float) * L::sfloats;
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_surface_t<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_surface_t<N>, dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream, c->K,
                     c->fgeo, c->Fmask, c->vmapP, c->EEL_ids4, c->EEL_vals, c->L0_ids, c->L0_vals,
                     c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

hipError_t launch_volume_tiled(bbw_ctx *c) {
  switch (c->N) {
    case 1: return volume_t_n<1>(c);
    case 2: return volume_t_n<2>(c);
    case 3: return volume_t_n<3>(c);
    case 4: return volume_t_n<4>(c);
    case 5: return volume_t_n<5>(c);
    case 6: return volume_t_n<6>(c);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_surface_tiled(bbw_ctx *c) {
  switch (c->N) {
    case 1: return surface_t_n<1>(c);
    case 2: return surface_t_n<2>(c);
    case 3: return surface_t_n<3>(c);
    case 4: return surface_t_n<4>(c);
    case 5: return surface_t_n<5>(c);
    case 6: return surface_t_n<6>(c);
  }
  return hipErrorInvalidValue;
}

// Elastic WADG update on MFMA (rk_update_elas, elas_cpp/okl/ElasKernelsWADG.okl:286-433), per
// 16-element tile: the 9 fields go through Vq in three groups of three, whose cubature-point
// maps are independent (velocities x 1/rho; normal stresses x (lambda, mu); shear stresses
// x mu), then back through Pq; LSRK stage with the source.
template <int N>
struct ElasUpdLayout {
  static constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq, TB = 16, NpNf = 9 * Np;
  static constexpr int NQT = (Nq + 15) / 16, NQP = 16 * NQT, NPT = (Np + 15) / 16;
  // KTP: K-steps over the cubature points (S is zero past Nq, so the all-zero steps are skipped)
  static constexpr int KTV = (Np + 3) / 4, KTP = (Nq + 3) / 4, QTW = (NQT + 3) / 4;
  static constexpr int KBV = (KTV + 3) / 4, KBP = (KTP + 3) / 4;
  static constexpr int RS0 = 9 * Np > 8 * Np + 4 * KTV ? 9 * Np : 8 * Np + 4 * KTV;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  // rhs rows [TB][RS] | S [3 fields][NQP][16] | source slots [TB] | one scratch word
  static constexpr size_t floats = (size_t)TB * RS + 3 * NQP * 16 + TB + 1;
};

template <int N>
__global__ __launch_bounds__(256, N <= 4 ? 4 : N == 5 ? 3 : 2) void k_update_elas_mfma(
    int K, FusedOps o, const float *__restrict__ rhoinvq, const float *__restrict__ lambdaq,
    const float *__restrict__ muq, SrcArgs src, const float *ftime_dev, float ftime, float fa,
    float fb, float fdt, const float *__restrict__ rhs, float *__restrict__ res, float *__restrict__ Q) {
  using L = ElasUpdLayout<N>;
  constexpr int Np = L::Np, Nq = L::Nq, TB = L::TB, NpNf = L::NpNf, RS = L::RS, NQP = L::NQP;
  constexpr int KTV = L::KTV, KTP = L::KTP, KBV = L::KBV, KBP = L::KBP;
  constexpr unsigned ROW = NpNf * 4u;
  extern __shared__ float lds[];
  float *sR = lds;                         // [TB][RS]
  float *S = sR + TB * RS;                 // [3][NQP][16]
  int *sSlot = (int *)(S + 3 * NQP * 16);  // [TB]
  float *sScr = (float *)(sSlot + TB);     // written by lanes past a row, never read
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * NpNf;
  const unsigned tbytes = (unsigned)nk * ROW, qbytes = (unsigned)nk * Nq * 4u;
  const brsrc rRh = mkrsrc(rhs + base, tbytes), rRs = mkrsrc(res + base, tbytes);
  const brsrc rQ = mkrsrc(Q + base, tbytes);
  const brsrc rRho = mkrsrc(rhoinvq + (size_t)k0 * Nq, qbytes), rLam = mkrsrc(lambdaq + (size_t)k0 * Nq, qbytes);
  const brsrc rMu = mkrsrc(muq + (size_t)k0 * Nq, qbytes);
  const brsrc rAV = mkrsrc(o.VqBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rAP = mkrsrc(o.PqBlk, (unsigned)L::NPT * KBP * 1024u);
  static_assert(TB == 16, "the per-tile source table is built for 16-element tiles");
  const int2 stile = src_tile(src, k0, nk);
  const bool has_src = stile.y != 0;

  // ---- 1: rhs rows, materials at the cubature points (MFMA output layout), Vq and Pq blocks;
  // then the rows -> LDS
  TileRows<NpNf, TB> rows;
  rows.load(rRh, ROW, tid);
  float rho[L::QTW][4], lam[L::QTW][4], mu[L::QTW][4];
  f4 aV[L::QTW][KBV], aP[KBP];
  const unsigned co = (unsigned)(col * Nq + 4 * kr) * 4u;
  // unconditional, clamped to the last tile (as in the acoustic update)
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t < L::NQT ? wave + 4 * t : L::NQT - 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      rho[t][r] = bld(rRho, co + (qt * 16 + r) * 4u);
      lam[t][r] = bld(rLam, co + (qt * 16 + r) * 4u);
      mu[t][r] = bld(rMu, co + (qt * 16 + r) * 4u);
    }
#pragma unroll
    for (int kb = 0; kb < KBV; ++kb) aV[t][kb] = bld_blk<KTV>(rAV, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
  }
  {
    const int mt0 = wave < L::NPT ? wave : L::NPT - 1;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<KTP>(rAP, lane * 16u, (unsigned)(mt0 * KBP + kb) * 1024u, kb);
  }
  rows.template store<RS>(sR, sScr, tid);
  constexpr int KPAD = L::RS0 - NpNf;
  static_assert(TB * KPAD <= 256, "one padding word per thread");
  if (KPAD > 0 && tid < TB * KPAD) sR[(tid / KPAD) * RS + NpNf + tid % KPAD] = 0.f;
  if (has_src && tid < TB) sSlot[tid] = src_slot_tile(src, stile, k0, tid);
  __syncthreads();

#pragma unroll
  for (int g = 0; g < 3; ++g) {
    // ---- 2: S_b = map_g(Vq . rhs_{3g+b}) (M = cubature points, N = elements, K = nodes)
    {
      const float *Bp = sR + col * RS + 3 * g * Np + kr;
#pragma unroll
      for (int t = 0; t < L::QTW; ++t) {
        const int qt = wave + 4 * t;
        if (qt >= L::NQT) break;
        f4 acc[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBV; ++kb)
#pragma unroll
          for (int sb = 0; sb < 4; ++sb) {
            const int kt = 4 * kb + sb;
            if (kt < KTV) {
#pragma unroll
              for (int b = 0; b < 3; ++b)
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aV[t][kb][sb], Bp[b * Np + kt * 4], acc[b], 0, 0, 0);
            }
          }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = qt * 16 + 4 * kr + r;
          float s0, s1, s2;
          if (g == 0) {
            s0 = rho[t][r] * acc[0][r];
            s1 = rho[t][r] * acc[1][r];
            s2 = rho[t][r] * acc[2][r];
          } else if (g == 1) {
            const float A = 2.f * mu[t][r] + lam[t][r], l = lam[t][r];
            s0 = A * acc[0][r] + l * acc[1][r] + l * acc[2][r];
            s1 = A * acc[1][r] + l * acc[0][r] + l * acc[2][r];
            s2 = A * acc[2][r] + l * acc[0][r] + l * acc[1][r];
          } else {
            s0 = mu[t][r] * acc[0][r];
            s1 = mu[t][r] * acc[1][r];
            s2 = mu[t][r] * acc[2][r];
          }
          const bool in = q < Nq;
          S[q * 16 + col] = in ? s0 : 0.f;
          S[(NQP + q) * 16 + col] = in ? s1 : 0.f;
          S[(2 * NQP + q) * 16 + col] = in ? s2 : 0.f;
        }
      }
    }
    __syncthreads();
    // ---- 3: rhs_{3g+b} = Pq . S_b (M = nodes, K = cubature points), over the group's rows
    for (int mt = wave; mt < L::NPT; mt += 4) {
      if (mt != wave) {
#pragma unroll
        for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<KTP>(rAP, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
      }
      f4 acc[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
      const float *Bp = S + kr * 16 + col;
#pragma unroll
      for (int kb = 0; kb < KBP; ++kb)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < KTP) {
#pragma unroll
            for (int b = 0; b < 3; ++b)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[kb][sb], Bp[(b * NQP + kt * 4) * 16], acc[b], 0, 0, 0);
          }
        }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int b = 0; b < 3; ++b) sR[col * RS + (3 * g + b) * Np + n] = acc[b][r];
        }
      }
      if (mt != wave && wave < L::NPT) {
#pragma unroll
        for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<KTP>(rAP, lane * 16u, (unsigned)(wave * KBP + kb) * 1024u, kb);
      }
    }
    __syncthreads();
  }

  // ---- 4: LSRK stage: res = fa res + fdt (rhs + source); Q += fb res
  const float ft = ftime_dev ? *ftime_dev : ftime;
  lsrk_rows<NpNf, RS, TB, Np>(rRs, rQ, ROW, sR, sSlot, has_src, src, ft, fa, fb, fdt, tid, [&](int fr) {
    float wsel = 0.f;
#pragma unroll
    for (int f = 0; f < 9; ++f) wsel = fr == f ? src.w[f] : wsel;
    return wsel;
  });
}

template <int N>
static hipError_t update_elas_mfma_n(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                     float ft, const SrcArgs &s) {
  using L = ElasUpdLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if ((unsigned long long)c->K * 36 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_update_elas_mfma<N>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_update_elas_mfma<N>, dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream,
                     c->K, fused_ops(c), c->rhoinv, c->lambda, c->mu, s, ftd, ft, fa, fb, fdt, c->rhs,
                     c->res, c->Q[c->cur]);
  return hipGetLastError();
}

hipError_t launch_update_elas_mfma(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                   float ft) {
  SrcArgs s = ctx_src(c);
  switch (c->N) {
    case 1: return update_elas_mfma_n<1>(c, fa, fb, fdt, ftd, ft, s);
    case 2: return update_elas_mfma_n<2>(c, fa, fb, fdt, ftd, ft, s);
    case 3: return update_elas_mfma_n<3>(c, fa, fb, fdt, ftd, ft, s);
    case 4: return update_elas_mfma_n<4>(c, fa, fb, fdt, ftd, ft, s);
    case 5: return update_elas_mfma_n<5>(c, fa, fb, fdt, ftd, ft, s);
    case 6: return update_elas_mfma_n<6>(c, fa, fb, fdt, ftd, ft, s);
  }
  return hipErrorInvalidValue;
}

// Elastic surface kernel on 16-element tiles (rk_surface_elas,
// elas_cpp/okl/ElasKernelsWADG.okl:128-284, whose lift is a dense Np x 4Nfp product): central
// flux with tau_v / tau_s penalties (Q+ = 0 on the boundary) for the 9 fields in registers,
// then, three fields at a time, the BB lift LIFT = EEL blkdiag(L0) on MFMA as one product per
// face (M = nodes, N = 16 elements, K = face nodes), accumulated into rhs.
template <int N>
struct ElasSurfLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, TB = 16;
  static constexpr int RowF = 9 * Np, QT = TB * RowF;
  static constexpr unsigned QB = QT * 4u, QIT = (QB + 4095) / 4096;
  static constexpr int NG3 = 256 / NfpNf, EPT3 = (TB + NG3 - 1) / NG3;
  static constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS, EPTV = (TB + NG - 1) / NG;
  static constexpr int KF = 4 * ((Nfp + 3) / 4), KTL = KF / 4, KBL = (KTL + 3) / 4;
  static constexpr int NPT = (Np + 15) / 16;
  static constexpr int KB = NfpNf + 4, BS = TB + 1;  // flux rows (zero tail) and row stride
  static constexpr int GB = 3 * KB * BS;             // a group's flux, B operand layout
  static constexpr int GO = 3 * TB * Np;             // a group's lifted values [3][TB][Np]
  static constexpr int QTP = (int)QIT * 1024;        // the Q tile padded to whole 4 KB rounds
  static constexpr int W = QTP > GB + GO ? QTP : GB + GO;
  static constexpr size_t floats = (size_t)W + TB * 16;
  static_assert(3 * (KB - NfpNf) * BS <= 256, "one tail word per thread");
};

template <int N>
__global__ __launch_bounds__(256) void k_surface_elas_t(int K, const float *__restrict__ fgeo,
                                                        const int *__restrict__ Fmask,
                                                        const int *__restrict__ vmapP,
                                                        const float *__restrict__ LiftBlk,
                                                        float tau_v, float tau_s,
                                                        const float *__restrict__ Q, float *__restrict__ rhs) {
  using L = ElasSurfLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, TB = L::TB, RowF = L::RowF;
  constexpr int KB = L::KB, BS = L::BS, KTL = L::KTL, KBL = L::KBL, NPT = L::NPT;
  extern __shared__ float lds[];
  float *sQ = lds;                 // [TB][9][Np]                  (phases 1-2)
  float *sB = lds;                 // [3][KB][BS] group flux       (group phases)
  float *sO = lds + L::GB;         // [3][TB][Np] lifted values    (group phases)
  float *sF = lds + L::W;          // [TB][16]
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const brsrc rO = mkrsrc(rhs + (size_t)k0 * RowF, (unsigned)nk * RowF * 4u);
  const brsrc rT = mkrsrc(Q, (unsigned)K * RowF * 4u);
  // ---- 1: the thread's face node and vmapP entries, the Q tile, fgeo, then the neighbour
  // traces (9 fields).  Every load is issued before the tile's LDS stores, which are
  // unconditional (the region holds whole 4 KB rounds): a conditional store lets the compiler
  // sink its load and serialise the rounds.
  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf;
  const bool on3 = g3 < L::NG3;
  const int f3 = m3 / Nfp;
  const int fid3 = on3 ? Fmask[m3] : 0;
  int vm[L::EPT3];
  {
    const brsrc rv = mkrsrc(vmapP + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) vm[i] = bldi(rv, (unsigned)((g3 + L::NG3 * i) * NfpNf + m3) * 4u);
  }
  float4 qreg[L::QIT];
  {
    const brsrc rQ = mkrsrc(Q + (size_t)k0 * RowF, (unsigned)nk * RowF * 4u);
#pragma unroll
    for (unsigned it = 0; it < L::QIT; ++it) qreg[it] = bld4(rQ, tid * 16u + it * 4096u);
  }
  const float fgv = bld(mkrsrc(fgeo + (size_t)k0 * 16, (unsigned)nk * 64u), tid * 4u);
  float fl[L::EPT3][9];
#pragma unroll
  for (int i = 0; i < L::EPT3; ++i)
#pragma unroll
    for (int c = 0; c < 9; ++c) fl[i][c] = bld(rT, (unsigned)(vm[i] + c * Np) * 4u);
  {
    static_assert(L::QIT * 1024 <= L::W, "the tile region holds whole rounds");
    float4 *sQ4 = (float4 *)sQ;
#pragma unroll
    for (unsigned it = 0; it < L::QIT; ++it) sQ4[tid + it * 256u] = qreg[it];
  }
  sF[tid] = fgv;
  __syncthreads();
  // ---- 2: flux per face node (into the same registers)
  if (on3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) {
        const float *q = sQ + e * RowF + fid3;
        const float *F = sF + 16 * e + 4 * f3;
        const float Fs = .5f * F[0], nx = F[1], ny = F[2], nz = F[3];
        const bool isB = vm[i] == fid3 + (k0 + e) * RowF;
        float dQ[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) dQ[c] = (isB ? 0.f : fl[i][c]) - q[c * Np];
        float fc[9];
        fc[0] = dQ[3] * nx + dQ[8] * ny + dQ[7] * nz;
        fc[1] = dQ[8] * nx + dQ[4] * ny + dQ[6] * nz;
        fc[2] = dQ[7] * nx + dQ[6] * ny + dQ[5] * nz;
        fc[3] = dQ[0] * nx;
        fc[4] = dQ[1] * ny;
        fc[5] = dQ[2] * nz;
        fc[6] = dQ[2] * ny + dQ[1] * nz;
        fc[7] = dQ[2] * nx + dQ[0] * nz;
        fc[8] = dQ[1] * nx + dQ[0] * ny;
        fl[i][0] = Fs * (fc[0] + tau_v * (fc[3] * nx + fc[8] * ny + fc[7] * nz));
        fl[i][1] = Fs * (fc[1] + tau_v * (fc[8] * nx + fc[4] * ny + fc[6] * nz));
        fl[i][2] = Fs * (fc[2] + tau_v * (fc[7] * nx + fc[6] * ny + fc[5] * nz));
        fl[i][3] = Fs * (fc[3] + tau_s * (fc[0] * nx));
        fl[i][4] = Fs * (fc[4] + tau_s * (fc[1] * ny));
        fl[i][5] = Fs * (fc[5] + tau_s * (fc[2] * nz));
        fl[i][6] = Fs * (fc[6] + tau_s * (fc[2] * ny + fc[1] * nz));
        fl[i][7] = Fs * (fc[7] + tau_s * (fc[2] * nx + fc[0] * nz));
        fl[i][8] = Fs * (fc[8] + tau_s * (fc[1] * nx + fc[0] * ny));
      }
    }
  }
  __syncthreads();
  const int gv = tid / L::GS, nv = tid - gv * L::GS;
  const bool onv = gv < L::NG && nv < Np;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  // lane-ordered LIFT_f blocks of the wave's first node tile (every group uses them)
  const brsrc rAL = mkrsrc(LiftBlk, 4u * NPT * KBL * 1024u);
  f4 aL[4][KBL];
  auto load_lift = [&](int mt) {
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int b = 0; b < KBL; ++b) aL[f][b] = bld_blk<KTL>(rAL, lane * 16u, (unsigned)((f * NPT + mt) * KBL + b) * 1024u, b);
  };
  load_lift(wave < NPT ? wave : NPT - 1);   // unconditional: see the MFMA updates
  // unrolled: the group selects fl[.][3 g + b] at compile time (a run-time index would move
  // the flux registers to scratch memory)
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    // ---- 3: the group's flux -> LDS as MFMA B operands (rows = face nodes, columns =
    // elements), with zero rows past the last face
    if (on3) {
#pragma unroll
      for (int i = 0; i < L::EPT3; ++i) {
        const int e = g3 + L::NG3 * i;
        if (L::NG3 * L::EPT3 == TB || e < TB) {
#pragma unroll
          for (int b = 0; b < 3; ++b) sB[(b * KB + m3) * BS + e] = fl[i][3 * g + b];
        }
      }
    }
    if (tid < 3 * (KB - NfpNf) * BS) {
      const int b = tid / ((KB - NfpNf) * BS), j = tid - b * ((KB - NfpNf) * BS);
      sB[(b * KB + NfpNf) * BS + j] = 0.f;
    }
    __syncthreads();
    // ---- 4: lift on MFMA: rhs_c += sum_f LIFT_f flux_{c,f} for the group's three fields c
    {
      const float *Bp = sB + kr * BS + col;
      for (int mt = wave; mt < NPT; mt += 4) {
        // the registers hold the wave's first tile, except after a wave's second tile
        if (mt != wave || (g > 0 && NPT > wave + 4)) load_lift(mt);
        f4 acc[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          acc[b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int kb = 0; kb < KBL; ++kb)
#pragma unroll
              for (int sb = 0; sb < 4; ++sb) {
                const int kt = 4 * kb + sb;
                if (kt < KTL)
                  acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aL[f][kb][sb], Bp[(b * KB + f * Nfp + 4 * kt) * BS],
                                                                acc[b], 0, 0, 0);
              }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = mt * 16 + 4 * kr + r;
          if (n < Np) {
#pragma unroll
            for (int b = 0; b < 3; ++b) sO[(b * TB + col) * Np + n] = acc[b][r];
          }
        }
      }
    }
    __syncthreads();
    // ---- 5: the group's rhs rows += lifted values (lanes over nodes)
    if (onv) {
      float ro[L::EPTV][3];
#pragma unroll
      for (int i = 0; i < L::EPTV; ++i)
#pragma unroll
        for (int b = 0; b < 3; ++b)
          ro[i][b] = bld(rO, (unsigned)((gv + L::NG * i) * RowF + (3 * g + b) * Np + nv) * 4u);
#pragma unroll
      for (int i = 0; i < L::EPTV; ++i) {
        const int e = gv + L::NG * i;
#pragma unroll
        for (int b = 0; b < 3; ++b)
          bst(rO, ro[i][b] + sO[(b * TB + e) * Np + nv], (unsigned)(e * RowF + (3 * g + b) * Np + nv) * 4u);
      }
    }
  }
}

template <int N>
static hipError_t surface_elas_t_n(bbw_ctx *c) {
  using L = ElasSurfLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if ((unsigned long long)c->K * 36 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_surface_elas_t<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_surface_elas_t<N>, dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream, c->K,
                     c->fgeo, c->Fmask, c->vmapP, c->LiftBlk, c->tau_v, c->tau_s, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

hipError_t launch_surface_elas_tiled(bbw_ctx *c) {
  switch (c->N) {
    case 1: return surface_elas_t_n<1>(c);
    case 2: return surface_elas_t_n<2>(c);
    case 3: return surface_elas_t_n<3>(c);
    case 4: return surface_elas_t_n<4>(c);
    case 5: return surface_elas_t_n<5>(c);
    case 6: return surface_elas_t_n<6>(c);
  }
  return hipErrorInvalidValue;
}

// Elastic volume kernel on 16-element tiles (rk_volume_elas, elas_cpp/okl/ElasKernelsWADG.okl
// :25-124, with the BB derivative operators) in the factored form of the fused acoustic stage.
// D_m u[a] = sum_j a_j u[a - e_j + e_m], and the nine outputs (div sigma and the six strain
// rates; the update applies the stiffness) are linear in the D_m u, so
// each output is the elevation out[a] = sum_j a_j F[a - e_j] of values F formed at the
// degree-(N-1) nodes b from the shifted coefficients u[b + e_m] (step 1, 36 scalar LDS reads per
// b), followed by 12 LDS reads per node (step 2), against 48 (16 per operator entry of the
// unfactored 4-nnz rows) before.
template <int N>
struct ElasVolLayout {
  static constexpr int Np = Dims<N>::Np, TB = 16, RowF = 9 * Np, QT = TB * RowF;
  static constexpr unsigned QB = QT * 4u, QIT = (QB + 4095) / 4096;
  static constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS, EPTV = (TB + NG - 1) / NG;
  static constexpr int Np1 = N * (N + 1) * (N + 2) / 6;
  static constexpr int GS1 = Np1 <= 32 ? 32 : 64, NG1 = 256 / GS1, EPT1 = TB / NG1;
  static_assert(Np1 <= 64 && NG1 * EPT1 == TB, "step 1 covers the tile");
  // Q tile [TB][9][Np] (padded to whole 4 KB rounds), then in the same region F as
  // [TB][Np1] float4 | [TB][Np1] float4 | [TB][Np1] float; vgeo in a 256-word slot
  static constexpr int QTP = (int)QIT * 1024;
  static_assert(9 * TB * Np1 <= QTP, "F fits the tile region");
  static constexpr size_t floats = (size_t)QTP + 256;
};

template <int N>
__global__ __launch_bounds__(256) void k_volume_elas_t(int K, const float *__restrict__ vgeo,
                                                       const int4 *__restrict__ Sids,
                                                       const int4 *__restrict__ Eids,
                                                       const float4 *__restrict__ Evals,
                                                       const float *__restrict__ Q, float *__restrict__ rhs) {
  using L = ElasVolLayout<N>;
  constexpr int Np = L::Np, TB = L::TB, RowF = L::RowF, Np1 = L::Np1;
  extern __shared__ float lds[];
  float *sQ = lds;                             // [TB][9][Np]     (step 1)
  float4 *sFa = (float4 *)lds;                 // [TB][Np1] F0-3  (step 2)
  float4 *sFb = sFa + TB * Np1;                // [TB][Np1] F4-7
  float *sFc = (float *)(sFb + TB * Np1);      // [TB][Np1] F8
  float *sG = lds + L::QTP;                    // [TB][9] in a 256-word slot
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const int gv = tid / L::GS, nv = tid - gv * L::GS;
  const bool onv = gv < L::NG && nv < Np;
  const int g1 = tid / L::GS1, n1 = tid - g1 * L::GS1;
  const bool on1 = n1 < Np1;
  // the thread's rows of the shift and elevation tables, loaded with the tile (lanes past the
  // nodes read row 0 and do not use it)
  const int4 S = Sids[on1 ? n1 : 0];
  const int4 E = Eids[nv < Np ? nv : 0];
  const float4 ev = Evals[nv < Np ? nv : 0];
  {
    // every load before the LDS stores, which are unconditional (whole rounds; words past the
    // tile and past vgeo's TB x 9 are written but never read)
    const brsrc rQ = mkrsrc(Q + (size_t)k0 * RowF, (unsigned)nk * RowF * 4u);
    float4 qreg[L::QIT];
#pragma unroll
    for (unsigned it = 0; it < L::QIT; ++it) qreg[it] = bld4(rQ, tid * 16u + it * 4096u);
    const float gv0 = bld(mkrsrc(vgeo + (size_t)k0 * 9, (unsigned)nk * 36u), tid * 4u);
    float4 *sQ4 = (float4 *)sQ;
#pragma unroll
    for (unsigned it = 0; it < L::QIT; ++it) sQ4[tid + it * 256u] = qreg[it];
    sG[tid] = gv0;
  }
  __syncthreads();
  // ---- step 1: F at the degree-(N-1) nodes b from u[b + e_m] (m = 1..4), into registers
  float F[L::EPT1][9];
  if (on1) {
#pragma unroll
    for (int i = 0; i < L::EPT1; ++i) {
      const int e = g1 + L::NG1 * i;
      const float *q = sQ + e * RowF;
      const float *G = sG + 9 * e;
      float Qx[9], Qy[9], Qz[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const float *qc = q + c * Np;
        const float s0 = qc[S.x], s1 = qc[S.y], s2 = qc[S.z], s3 = qc[S.w];
        const float dr = .5f * (s1 - s0), ds = .5f * (s2 - s0), dt = .5f * (s3 - s0);
        Qx[c] = G[0] * dr + G[3] * ds + G[6] * dt;
        Qy[c] = G[1] * dr + G[4] * ds + G[7] * dt;
        Qz[c] = G[2] * dr + G[5] * ds + G[8] * dt;
      }
      F[i][0] = Qx[3] + Qy[8] + Qz[7];
      F[i][1] = Qx[8] + Qy[4] + Qz[6];
      F[i][2] = Qx[7] + Qy[6] + Qz[5];
      F[i][3] = Qx[0];
      F[i][4] = Qy[1];
      F[i][5] = Qz[2];
      F[i][6] = Qy[2] + Qz[1];
      F[i][7] = Qx[2] + Qz[0];
      F[i][8] = Qx[1] + Qy[0];
    }
  }
  __syncthreads();
  if (on1) {
#pragma unroll
    for (int i = 0; i < L::EPT1; ++i) {
      const int e = g1 + L::NG1 * i;
      sFa[e * Np1 + n1] = make_float4(F[i][0], F[i][1], F[i][2], F[i][3]);
      sFb[e * Np1 + n1] = make_float4(F[i][4], F[i][5], F[i][6], F[i][7]);
      sFc[e * Np1 + n1] = F[i][8];
    }
  }
  __syncthreads();
  if (!onv) return;
  // ---- step 2: out[a] = sum_j a_j F[a - e_j]; rows out (lanes over nodes)
  const brsrc rO = mkrsrc(rhs + (size_t)k0 * RowF, (unsigned)nk * RowF * 4u);
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i) {
    const int e = gv + L::NG * i;
    const float4 *Fa = sFa + e * Np1, *Fb = sFb + e * Np1;
    const float *Fc = sFc + e * Np1;
    const float4 a1 = Fa[E.x], a2 = Fa[E.y], a3 = Fa[E.z], a4 = Fa[E.w];
    const float4 b1 = Fb[E.x], b2 = Fb[E.y], b3 = Fb[E.z], b4 = Fb[E.w];
    const float c1 = Fc[E.x], c2 = Fc[E.y], c3 = Fc[E.z], c4 = Fc[E.w];
    const unsigned o = (unsigned)(e * RowF + nv) * 4u;
    bst(rO, ev.x * a1.x + ev.y * a2.x + ev.z * a3.x + ev.w * a4.x, o);
    bst(rO, ev.x * a1.y + ev.y * a2.y + ev.z * a3.y + ev.w * a4.y, o + 1 * Np * 4u);
    bst(rO, ev.x * a1.z + ev.y * a2.z + ev.z * a3.z + ev.w * a4.z, o + 2 * Np * 4u);
    bst(rO, ev.x * a1.w + ev.y * a2.w + ev.z * a3.w + ev.w * a4.w, o + 3 * Np * 4u);
    bst(rO, ev.x * b1.x + ev.y * b2.x + ev.z * b3.x + ev.w * b4.x, o + 4 * Np * 4u);
    bst(rO, ev.x * b1.y + ev.y * b2.y + ev.z * b3.y + ev.w * b4.y, o + 5 * Np * 4u);
    bst(rO, ev.x * b1.z + ev.y * b2.z + ev.z * b3.z + ev.w * b4.z, o + 6 * Np * 4u);
    bst(rO, ev.x * b1.w + ev.y * b2.w + ev.z * b3.w + ev.w * b4.w, o + 7 * Np * 4u);
    bst(rO, ev.x * c1 + ev.y * c2 + ev.z * c3 + ev.w * c4, o + 8 * Np * 4u);
  }
}

template <int N>
static hipError_t volume_elas_t_n(bbw_ctx *c) {
  using L = ElasVolLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if ((unsigned long long)c->K * 36 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_volume_elas_t<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_volume_elas_t<N>, dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream, c->K,
                     c->vgeo, c->Sids, c->Eids, c->Evals, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

hipError_t launch_volume_elas_tiled(bbw_ctx *c) {
  switch (c->N) {
    case 1: return volume_elas_t_n<1>(c);
    case 2: return volume_elas_t_n<2>(c);
    case 3: return volume_elas_t_n<3>(c);
    case 4: return volume_elas_t_n<4>(c);
    case 5: return volume_elas_t_n<5>(c);
    case 6: return volume_elas_t_n<6>(c);
  }
  return hipErrorInvalidValue;
}

// Exact Bernstein-Bezier WADG pressure update (rk_update_BBWADG,
// cpp/okl/WaveKernelsBBWADG.okl:14-321): rhs_p <- P_N(c2 rhs_p), c2 in degree-N BB coefficients.
// One workgroup per 16-element tile:
//   1  rows C(N,a) c2_a and C(N,b) p_b -> LDS; the degree-2N product tile f [g][element] zeroed
//   2  f_g = sum_{a+b=g} (C(N,a) c2_a)(C(N,b) p_b), four elements per wave with lanes over b: for
//      each a the lanes' targets g = a + b are distinct, so the LDS accumulation is race free
//      (the 1/C(2N,g) scaling is folded into the projection)
//   3  r = P' f on MFMA 16x16x4 (M = nodes, N = 16 elements, K = degree-2N coefficients),
//      P' = sum_j c_j E_{N-j}^N (E_{N-j}^2N)^T diag(1/C(2N,g)) as lane-ordered blocks
//   4  r -> rhs_p rows through LDS
template <int N>
struct BBLayout {
  static constexpr int Np = Dims<N>::Np, N2p = (2 * N + 1) * (2 * N + 2) * (2 * N + 3) / 6, TB = 16;
  static constexpr int KT = (N2p + 3) / 4, K2 = 4 * KT, KB = (KT + 3) / 4, NPT = (Np + 15) / 16;
  static constexpr int FS = TB + 1;  // f row stride: element columns + 1 (bank spread)
  static constexpr int PR = (TB * Np + 255) / 256;  // 256-thread rounds of the row loads
  // sp [TB][Np] | sc [TB][Np] (each PR x 256 words) | f [K2][FS]
  static constexpr size_t floats = (size_t)2 * PR * 256 + (size_t)K2 * FS;
};

template <int N>
__global__ __launch_bounds__(256) void k_update_bbwadg(int K, const float *__restrict__ CN,
                                                       const float *__restrict__ PBlk,
                                                       const float *__restrict__ c2bb,
                                                       float *__restrict__ rhs) {
  using L = BBLayout<N>;
  constexpr int Np = L::Np, TB = L::TB, FS = L::FS, K2 = L::K2, KT = L::KT, KB = L::KB;
  extern __shared__ float lds[];
  float *sp = lds, *sc = sp + L::PR * 256, *f = sc + L::PR * 256;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  // ---- 1: all row loads issued before the LDS stores (whole rounds, unconditional; rows of
  // elements past the mesh and words past the tile read as zero through the range checks)
  {
    const brsrc rp = mkrsrc(rhs + (size_t)k0 * 4 * Np, (unsigned)nk * 16u * Np);
    const brsrc rc = mkrsrc(c2bb + (size_t)k0 * Np, (unsigned)nk * 4u * Np);
    float pv[L::PR], cv[L::PR], cn[L::PR];
#pragma unroll
    for (int r = 0; r < L::PR; ++r) {
      const int i = tid + 256 * r, e = i / Np, n = i - e * Np;
      pv[r] = bld(rp, (unsigned)(e * 4 * Np + n) * 4u);
      cv[r] = bld(rc, (unsigned)i * 4u);
      cn[r] = CN[n];
    }
    for (int i = tid; i < K2 * FS; i += 256) f[i] = 0.f;
#pragma unroll
    for (int r = 0; r < L::PR; ++r) {
      sp[tid + 256 * r] = cn[r] * pv[r];
      sc[tid + 256 * r] = cn[r] * cv[r];
    }
  }
  __syncthreads();
  // ---- 2: the wave's elements wave + 4 i (i < 4) together, lanes over b (and b + 64 when
  // Np > 64).  Multi-indices (i, j, k, l) in the reference order (l outer, k, j inner); the
  // degree-2N index of (j, k, l) is Np(2N) - Np(2N - l) + k (2N - l + 1) - k (k - 1) / 2 + j, so for
  // fixed (k_a, l_a) it steps by one with j_a.  The lanes' targets a + b are distinct for each a
  // and only this wave updates its element columns, so the read-add-write updates need no
  // atomics (LDS float atomics measured 12x slower).
  {
    constexpr int M = 2 * N, NpM = L::N2p;
    auto npd = [](int d) { return (d + 1) * (d + 2) * (d + 3) / 6; };
#pragma unroll
    for (int b0 = 0; b0 < Np; b0 += 64) {
      const int b = b0 + lane;
      const bool on = b < Np;
      // (j, k, l) of node b
      int lb = 0, kb = 0, rem = on ? b : 0;
      while (rem >= (N - lb + 1) * (N - lb + 2) / 2) { rem -= (N - lb + 1) * (N - lb + 2) / 2; ++lb; }
      while (rem >= N - lb - kb + 1) { rem -= N - lb - kb + 1; ++kb; }
      const int jb = rem;
      float pb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) pb[i] = on ? sp[(wave + 4 * i) * Np + b] : 0.f;
      int a = 0;
      for (int la = 0; la <= N; ++la) {
        const int l = la + lb;
        for (int ka = 0; ka <= N - la; ++ka) {
          const int k = ka + kb;
          const int g0 = NpM - npd(M - l) + k * (M - l + 1) - k * (k - 1) / 2 + jb;
          for (int ja = 0; ja <= N - la - ka; ++ja, ++a) {
            float *fg = f + (g0 + ja) * FS + wave;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              if (on) fg[4 * i] += sc[(wave + 4 * i) * Np + a] * pb[i];
            }
            // other lanes update this lane's next targets: keep each step's LDS reads after the
            // previous step's writes (the compiler sees one lane, whose own addresses differ)
            asm volatile("" ::: "memory");
          }
        }
      }
    }
  }
  __syncthreads();
  // ---- 3: r = P' f; rows -> sp (its last reads were before the barrier)
  {
    const brsrc rP = mkrsrc(PBlk, (unsigned)L::NPT * KB * 1024u);
    const int col = lane & 15, kr = lane >> 4;
    const float *Bp = f + kr * FS + col;
    for (int mt = wave; mt < L::NPT; mt += 4) {
      // four independent accumulation chains (K-steps 4 kb + sb), summed at the end
      f4 acc[4];
#pragma unroll
      for (int sb = 0; sb < 4; ++sb) acc[sb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
      for (int kb = 0; kb < KB; ++kb) {
        const f4 a4 = bld4_s(rP, lane * 16u, (unsigned)(mt * KB + kb) * 1024u);
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < KT) acc[sb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[sb], Bp[kt * 4 * FS], acc[sb], 0, 0, 0);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) sp[col * Np + n] = (acc[0][r] + acc[1][r]) + (acc[2][r] + acc[3][r]);
      }
    }
  }
  __syncthreads();
  // ---- 4
  for (int i = tid; i < nk * Np; i += 256) {
    const int e = i / Np, n = i - e * Np;
    rhs[(size_t)(k0 + e) * 4 * Np + n] = sp[i];
  }
}

template <int N>
static hipError_t update_bbwadg_n(bbw_ctx *c) {
  using L = BBLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_update_bbwadg<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_update_bbwadg<N>), dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream, c->K,
                     c->bb_CN, c->bb_PBlk, c->c2bb, c->rhs);
  return hipGetLastError();
}

hipError_t launch_update_bbwadg(bbw_ctx *c) {
  switch (c->N) {
    case 1: return update_bbwadg_n<1>(c);
    case 2: return update_bbwadg_n<2>(c);
    case 3: return update_bbwadg_n<3>(c);
    case 4: return update_bbwadg_n<4>(c);
    case 5: return update_bbwadg_n<5>(c);
    case 6: return update_bbwadg_n<6>(c);
  }
  return hipErrorInvalidValue;
}

FusedOps fused_ops(const bbw_ctx *c) {
  FusedOps o;
  o.Fmask = c->Fmask;
  o.Dids = c->Dids;
  o.Dvals = c->Dvals;
  o.EEL_ids4 = c->EEL_ids4;
  o.EEL_vals = c->EEL_vals;
  o.L0_ids = c->L0_ids;
  o.L0_vals = c->L0_vals;
  o.VqPad = c->VqPad;
  o.PqPad = c->PqPad;
  o.LiftPad = c->LiftPad;
  o.VqBlk = c->VqBlk;
  o.PqBlk = c->PqBlk;
  o.LiftBlk = c->LiftBlk;
  o.VqCBlk = c->VqCBlk;
  o.Sids = c->Sids;
  o.Eids = c->Eids;
  o.Evals = c->Evals;
  return o;
}

template <int N, bool P>
static hipError_t fused_mfma_n(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                               float fdt, const float *ftd, float ft, const SrcArgs &s,
                               float *yout, int store_res) {
  using L = FusedLayout<N, P>;
  const size_t sh = sizeof(float) * L::floats;
  const int grid = (c->K + L::TB - 1) / L::TB;
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused_mfma<N, P>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  FusedMesh m{c->K, c->vgeo, c->fgeo, c->vmapP, c->xcd_chunk};
  FusedShot shot{c->c2q, Qin, Qout, c->res, yout, ft, fdt, store_res};
  hipLaunchKernelGGL((k_fused_mfma<N, P>), dim3(grid), dim3(256), sh, c->stream, m, fused_ops(c),
                     shot, s, fa, fb, ftd);
  return hipGetLastError();
}

template <int N>
static hipError_t fused_mfma_sel(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                                 float fdt, const float *ftd, float ft, int projU,
                                 const SrcArgs &s, float *yout, int store_res) {
  return projU ? fused_mfma_n<N, true>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, yout, store_res)
               : fused_mfma_n<N, false>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, yout, store_res);
}

hipError_t launch_fused_mfma(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                             float fdt, const float *ftd, float ft, int projU, float *yout,
                             int store_res) {
  SrcArgs s = ctx_src(c);
  switch (c->N) {
    case 1: return fused_mfma_sel<1>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout, store_res);
    case 2: return fused_mfma_sel<2>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout, store_res);
    case 3: return fused_mfma_sel<3>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout, store_res);
    case 4: return fused_mfma_sel<4>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout, store_res);
    case 5: return fused_mfma_sel<5>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout, store_res);
    case 6: return fused_mfma_sel<6>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout, store_res);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
