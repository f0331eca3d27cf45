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
  const int k0 = blockIdx.x * L::TB, nk = min(L::TB, m.K - k0);
  if (ftime_dev) sh.ft = *ftime_dev;
  TileStage<N, PROJ_U> st;
  stage_load<N, PROJ_U>(st, m, sh.Qin, k0, nk, tid);
  int vm[L::EPT3];
  vm_load<N, PROJ_U>(vm, m, k0, nk, tid);
  stage_store<N, PROJ_U>(st, lds + L::OQ, lds + L::OG, tid);
  __syncthreads();
  tile_shot<N, PROJ_U>(m, o, sh, src, fa, fb, k0, nk, lds, lds + L::OQ, lds + L::OG, vm,
                       face_node<N, PROJ_U>(o, tid), tid, NoHook{}, NoHook{}, NoHook{});
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
  // rhs rows [TB][RS] | S [4 fields][NQP][16] | source slots [TB]
  static constexpr size_t floats = (size_t)TB * RS + 4 * NQP * 16 + TB;
};

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
  const bool has_src = src.nsrc > 0 && k0 + nk - 1 >= src.first && k0 <= src.last;

  // ---- 1: rhs tile -> LDS; res and Q rows, c2 and 1/J (MFMA output layout), Vq blocks
  float rreg[L::PF ? TB : 1], qreg[L::PF ? TB : 1];
  if constexpr (L::PF) {
#pragma unroll
    for (int e = 0; e < TB; ++e) {
      const float v = bld(rRh, tid * 4u + e * ROW);
      if (tid < NpNf) sR[e * RS + tid] = v;
      rreg[e] = bld(rRs, tid * 4u + e * ROW);
      qreg[e] = bld(rQ, tid * 4u + e * ROW);
    }
  } else {
    for (int rr = tid; rr < NpNf; rr += 256)
      for (int e = 0; e < TB; ++e) sR[e * RS + rr] = bld(rRh, rr * 4u + e * ROW);
  }
  constexpr int KPAD = L::RS0 - NpNf;
  static_assert(TB * KPAD <= 256, "one padding word per thread");
  if (KPAD > 0 && tid < TB * KPAD) sR[(tid / KPAD) * RS + NpNf + tid % KPAD] = 0.f;
  if (has_src && tid < TB) sSlot[tid] = tid < nk ? src_slot2(src, k0 + tid) : -1;
  float c2r[L::QTW][4], jr[L::QTW][4];
  f4 aV[L::QTW][KBV];
  const unsigned co = (unsigned)(col * Nq + 4 * kr) * 4u;
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t;
    if (qt < L::NQT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        c2r[t][r] = bld(rC, co + (qt * 16 + r) * 4u);
        jr[t][r] = Jq ? 1.f / bld(rJ, co + (qt * 16 + r) * 4u) : 1.f;
      }
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb) aV[t][kb] = bld4_s(rAV, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u);
    }
  }
  __syncthreads();

  // ---- 2: S_f = (Vq . rhs_f) x c2/J (p) or 1/J (M = cubature points, N = elements, K = nodes)
  f4 aP[KBP];
  if (wave < L::NPT) {
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld4_s(rAP, lane * 16u, (unsigned)(wave * KBP + kb) * 1024u);
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
      for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld4_s(rAP, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u);
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
    for (int rr = tid; rr < NpNf; rr += 256) {
      const int fr = rr / Np, nr = rr - fr * Np;
      for (int e = 0; e < nk; ++e) {
        float v = sR[e * RS + rr];
        if (has_src) {
          const int slot = sSlot[e];
          if (slot >= 0) v += w4(src, fr) * ft * src.coef[(size_t)slot * Np + nr];
        }
        const unsigned g = rr * 4u + e * ROW;
        const float x = fa * bld(rRs, g) + fdt * v;
        bst(rRs, x, g);
        bst(rQ, bld(rQ, g) + fb * x, g);
      }
    }
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

hipError_t launch_update_wadg_mfma(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                   float ft) {
  SrcArgs s;
  s.nsrc = c->nsrc;
  s.first = c->src_first;
  s.last = c->src_last;
  s.elems = c->src_elems;
  s.coef = c->src_coef;
  for (int f = 0; f < 9; ++f) s.w[f] = c->src_w[f];
  switch (c->N) {
    case 1: return update_wadg_mfma_n<1>(c, fa, fb, fdt, ftd, ft, s);
    case 2: return update_wadg_mfma_n<2>(c, fa, fb, fdt, ftd, ft, s);
    case 3: return update_wadg_mfma_n<3>(c, fa, fb, fdt, ftd, ft, s);
    case 4: return update_wadg_mfma_n<4>(c, fa, fb, fdt, ftd, ft, s);
    case 5: return update_wadg_mfma_n<5>(c, fa, fb, fdt, ftd, ft, s);
    case 6: return update_wadg_mfma_n<6>(c, fa, fb, fdt, ftd, ft, s);
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
  return o;
}

template <int N, bool P>
static hipError_t fused_mfma_n(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                               float fdt, const float *ftd, float ft, const SrcArgs &s,
                               float *yout) {
  using L = FusedLayout<N, P>;
  const size_t sh = sizeof(float) * L::floats;
  const int grid = (c->K + L::TB - 1) / L::TB;
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused_mfma<N, P>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  FusedMesh m{c->K, c->vgeo, c->fgeo, c->vmapP};
  FusedShot shot{c->c2q, Qin, Qout, c->res, yout, ft, fdt};
  hipLaunchKernelGGL((k_fused_mfma<N, P>), dim3(grid), dim3(256), sh, c->stream, m, fused_ops(c),
                     shot, s, fa, fb, ftd);
  return hipGetLastError();
}

template <int N>
static hipError_t fused_mfma_sel(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                                 float fdt, const float *ftd, float ft, int projU,
                                 const SrcArgs &s, float *yout) {
  return projU ? fused_mfma_n<N, true>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, yout)
               : fused_mfma_n<N, false>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, yout);
}

hipError_t launch_fused_mfma(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                             float fdt, const float *ftd, float ft, int projU, float *yout) {
  SrcArgs s;
  s.nsrc = c->nsrc;
  s.first = c->src_first;
  s.last = c->src_last;
  s.elems = c->src_elems;
  s.coef = c->src_coef;
  for (int f = 0; f < 9; ++f) s.w[f] = c->src_w[f];
  switch (c->N) {
    case 1: return fused_mfma_sel<1>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 2: return fused_mfma_sel<2>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 3: return fused_mfma_sel<3>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 4: return fused_mfma_sel<4>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 5: return fused_mfma_sel<5>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 6: return fused_mfma_sel<6>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
