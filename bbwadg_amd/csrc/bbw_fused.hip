// Fused acoustic stage for gfx950 with the WADG projection on FP32 MFMA.
//
// One workgroup (4 waves) owns a tile of TB = 16 consecutive elements:
//   1  Q_in tile, geometry -> LDS (flat coalesced); c2 at cubature prefetched into registers
//      in the MFMA output layout
//   2  contravariant velocity (Ur, Us, Ut)
//   3  volume terms (4-nnz barycentric operators) -> rhs      [operator row in registers,
//                                                               reused for the group's elements]
//   4  upwind flux at face nodes, neighbour traces from Q_in
//   5  L0 per face                                             [row in registers]
//   6  EEL lift accumulated into rhs                           [row in registers]
//   7  S = (Vq . rhs_p) * c2      MFMA 16x16x4, M = cubature points, N = 16 elements, K = Np
//   8  rhs_p = Pq . S             MFMA 16x16x4, M = nodes, N = 16 elements, K = Nq
//   9  LSRK: res = fa res + fdt rhs (+ source); Q_out = Q_in + fb res (flat coalesced)
// With PROJ_U the velocity fields go through Pq . Vq as well (N = 64 columns).
//
// v_mfma_f32_16x16x4f32 lane layout (checked on gfx950): A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k, D[4*(l/16) + r][l%16] in register r of lane l.
#include "bbw_internal.h"

namespace bbw {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int src_slot2(const SrcArgs &s, int k) {
  if (s.nsrc == 0 || k < s.first || k > s.last) return -1;
  int lo = 0, hi = s.nsrc - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = s.elems[mid];
    if (v == k) return mid;
    if (v < k) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

template <int N, bool PROJ_U>
struct FusedLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf;
  static constexpr int Nq = Dims<N>::Nq;
  static constexpr int TB = 16;
  static constexpr int RS = 4 * Np + 1;          // padded element stride of the rhs tile
  static constexpr int NQT = (Nq + 15) / 16, NQP = NQT * 16;
  static constexpr int NPT = (Np + 15) / 16;
  static constexpr int KTV = (Np + 3) / 4, KTP = NQP / 4;
  static constexpr int NP = PROJ_U ? 4 : 1;
  static constexpr int SX = NP * TB;             // S row length
  static constexpr int X0 = 4 * NfpNf > 3 * Np ? 4 * NfpNf : 3 * Np;
  static constexpr int XS = X0 > (NQP * NP) ? X0 : NQP * NP;   // per element work floats
  static constexpr int QTW = (NQT + 3) / 4;     // q tiles per wave
  static constexpr size_t floats = (size_t)TB * (4 * Np + RS + 9 + 16 + XS) + TB;  // + slots
};

template <int N, bool PROJ_U>
__global__ __launch_bounds__(256) void k_fused_mfma(
    int K, const float *__restrict__ vgeo, const float *__restrict__ fgeo,
    const int *__restrict__ vmapP, const int *__restrict__ Fmask, const int4 *__restrict__ Dids,
    const float4 *__restrict__ Dvals, const int *__restrict__ EEL_ids,
    const float *__restrict__ EEL_vals, const int *__restrict__ L0_ids,
    const float *__restrict__ L0_vals, const float *__restrict__ Vq, const float *__restrict__ Pq,
    const float *__restrict__ c2q, float fa, float fb, float fdt, SrcArgs src,
    const float *ftime_dev, float ftime, const float *__restrict__ Qin, float *__restrict__ res,
    float *__restrict__ Qout) {
  using L = FusedLayout<N, PROJ_U>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, Nq = L::Nq, TB = L::TB, RS = L::RS;
  constexpr int Nf = 4, EEL_nnz = Dims<N>::EEL_nnz, L0_nnz = Dims<N>::L0_nnz;
  constexpr int NP = L::NP, SX = L::SX, XS = L::XS;
  extern __shared__ float lds[];
  float *sQ = lds;                  // [TB][4][Np]
  float *sR = sQ + TB * Nf * Np;    // [TB][RS]
  float *sG = sR + TB * RS;         // [TB][9]
  float *sF = sG + TB * 9;          // [TB][16]
  float *sX = sF + TB * 16;         // [TB][XS] work area
  int *sSlot = (int *)(sX + TB * XS);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k0 = blockIdx.x * TB;
  const int nk = min(TB, K - k0);

  // ---- 1: tile loads; c2 prefetch in the MFMA output layout
  {
    const float *Qt = Qin + (size_t)k0 * Nf * Np;
    for (int i = tid; i < nk * Nf * Np; i += 256) sQ[i] = Qt[i];
    for (int i = tid; i < nk * 9; i += 256) sG[i] = vgeo[(size_t)k0 * 9 + i];
    for (int i = tid; i < nk * 16; i += 256) sF[i] = fgeo[(size_t)k0 * 16 + i];
    for (int i = nk * RS + tid; i < TB * RS; i += 256) sR[i] = 0.f;
    if (tid < TB) sSlot[tid] = tid < nk ? src_slot2(src, k0 + tid) : -1;
  }
  float c2r[L::QTW][4];
  {
    const int e = lane & 15;
#pragma unroll
    for (int t = 0; t < L::QTW; ++t) {
      const int qt = wave + 4 * t;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * (lane >> 4) + r;
        c2r[t][r] = (qt < L::NQT && q < Nq && e < nk) ? c2q[(size_t)(k0 + e) * Nq + q] : 0.f;
      }
    }
  }
  __syncthreads();

  // ---- 2: contravariant velocity into the work area
  for (int w = tid; w < nk * Np; w += 256) {
    const int e = w / Np, n = w - e * Np;
    const float *s = sQ + e * Nf * Np;
    const float *G = sG + 9 * e;
    float *u = sX + e * XS;
    const float uu = s[Np + n], vv = s[2 * Np + n], ww = s[3 * Np + n];
    u[n] = uu * G[0] + vv * G[1] + ww * G[2];
    u[Np + n] = uu * G[3] + vv * G[4] + ww * G[5];
    u[2 * Np + n] = uu * G[6] + vv * G[7] + ww * G[8];
  }
  __syncthreads();

  // ---- 3: volume terms (thread <-> node, loop over the group's elements)
  {
    constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS;
    const int g = tid / GS, n = tid % GS;
    if (g < NG && n < Np) {
      const float4 dv = Dvals[n];
      const int4 D1 = Dids[n], D2 = Dids[Np + n], D3 = Dids[2 * Np + n], D4 = Dids[3 * Np + n];
      for (int e = g; e < nk; e += NG) {
        const float *sp = sQ + e * Nf * Np;
        const float *sUr = sX + e * XS, *sUs = sUr + Np, *sUt = sUr + 2 * Np;
        const float *G = sG + 9 * e;
        const float p1 = dv.x * sp[D1.x] + dv.y * sp[D2.x] + dv.z * sp[D3.x] + dv.w * sp[D4.x];
        const float p2 = dv.x * sp[D1.y] + dv.y * sp[D2.y] + dv.z * sp[D3.y] + dv.w * sp[D4.y];
        const float p3 = dv.x * sp[D1.z] + dv.y * sp[D2.z] + dv.z * sp[D3.z] + dv.w * sp[D4.z];
        const float p4 = dv.x * sp[D1.w] + dv.y * sp[D2.w] + dv.z * sp[D3.w] + dv.w * sp[D4.w];
        float dU1 = dv.x * (sUr[D1.y] + sUs[D1.z] + sUt[D1.w]);
        dU1 += dv.y * (sUr[D2.y] + sUs[D2.z] + sUt[D2.w]);
        dU1 += dv.z * (sUr[D3.y] + sUs[D3.z] + sUt[D3.w]);
        dU1 += dv.w * (sUr[D4.y] + sUs[D4.z] + sUt[D4.w]);
        float dU2 = dv.x * (sUr[D1.x] + sUs[D1.x] + sUt[D1.x]);
        dU2 += dv.y * (sUr[D2.x] + sUs[D2.x] + sUt[D2.x]);
        dU2 += dv.z * (sUr[D3.x] + sUs[D3.x] + sUt[D3.x]);
        dU2 += dv.w * (sUr[D4.x] + sUs[D4.x] + sUt[D4.x]);
        const float dpdr = .5f * (p2 - p1), dpds = .5f * (p3 - p1), dpdt = .5f * (p4 - p1);
        float *r = sR + e * RS;
        r[n] = -.5f * (dU1 - dU2);
        r[Np + n] = -(G[0] * dpdr + G[3] * dpds + G[6] * dpdt);
        r[2 * Np + n] = -(G[1] * dpdr + G[4] * dpds + G[7] * dpdt);
        r[3 * Np + n] = -(G[2] * dpdr + G[5] * dpds + G[8] * dpdt);
      }
    }
  }
  __syncthreads();

  // ---- 4: upwind flux at face nodes (own trace from LDS, neighbour trace from Q_in)
  for (int w = tid; w < nk * NfpNf; w += 256) {
    const int e = w / NfpNf, n = w - e * NfpNf, f = n / Nfp, k = k0 + e;
    const float *F = sF + 16 * e + 4 * f;
    const float *s = sQ + e * Nf * Np;
    const int fid = Fmask[n];
    const int idM = fid + k * Np * Nf;
    const int idP = vmapP[(size_t)k * NfpNf + n];
    const float pM = s[fid], uM = s[Np + fid], vM = s[2 * Np + fid], wM = s[3 * Np + fid];
    float pjump, Unjump;
    if (idM == idP) {
      pjump = -2.f * pM;
      Unjump = 0.f;
    } else {
      const float pP = Qin[idP], uP = Qin[idP + Np], vP = Qin[idP + 2 * Np], wP = Qin[idP + 3 * Np];
      pjump = pP - pM;
      Unjump = (uP - uM) * F[1] + (vP - vM) * F[2] + (wP - wM) * F[3];
    }
    float *X = sX + e * XS;
    X[n] = .5f * (pjump - Unjump) * F[0];
    X[NfpNf + n] = .5f * (Unjump - pjump) * F[0];
  }
  __syncthreads();

  // ---- 5: L0 per face (thread <-> face node)
  {
    constexpr int NG = 256 / NfpNf;
    const int g = tid / NfpNf, m = tid % NfpNf;
    if (g < NG) {
      const int f = m / Nfp, nt = m - f * Nfp;
      float Lv[L0_nnz];
      int Li[L0_nnz];
#pragma unroll
      for (int j = 0; j < L0_nnz; ++j) {
        Lv[j] = L0_vals[nt + j * Nfp];
        Li[j] = L0_ids[nt + j * Nfp] + f * Nfp;
      }
      for (int e = g; e < nk; e += NG) {
        float *X = sX + e * XS;
        float v1 = 0.f, v2 = 0.f;
#pragma unroll
        for (int j = 0; j < L0_nnz; ++j) {
          v1 += Lv[j] * X[Li[j]];
          v2 += Lv[j] * X[NfpNf + Li[j]];
        }
        X[2 * NfpNf + m] = v1;
        X[3 * NfpNf + m] = v2;
      }
    }
  }
  __syncthreads();

  // ---- 6: EEL lift into rhs (thread <-> node)
  {
    constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS;
    const int g = tid / GS, n = tid % GS;
    if (g < NG && n < Np) {
      float Ev[EEL_nnz];
      int Ei[EEL_nnz];
#pragma unroll
      for (int j = 0; j < EEL_nnz; ++j) {
        Ev[j] = EEL_vals[n + j * Np];
        Ei[j] = EEL_ids[n + j * Np];
      }
      for (int e = g; e < nk; e += NG) {
        const float *pt = sX + e * XS + 2 * NfpNf, *ut = pt + NfpNf, *F = sF + 16 * e;
        float *r = sR + e * RS;
        float v1 = r[n], v2 = r[Np + n], v3 = r[2 * Np + n], v4 = r[3 * Np + n];
#pragma unroll
        for (int j = 0; j < EEL_nnz; ++j) {
          const int col = Ei[j];
          const int fc = col / Nfp;
          const float Uf = Ev[j] * ut[col];
          v1 += Ev[j] * pt[col];
          v2 += Uf * F[4 * fc + 1];
          v3 += Uf * F[4 * fc + 2];
          v4 += Uf * F[4 * fc + 3];
        }
        r[n] = v1;
        r[Np + n] = v2;
        r[2 * Np + n] = v3;
        r[3 * Np + n] = v4;
      }
    }
  }
  __syncthreads();

  // ---- 7: S = (Vq . rhs) x c2 on MFMA; S is [NQP][SX] in the work area
  float *S = sX;
  {
    const int col = lane & 15, kr = lane >> 4;
#pragma unroll
    for (int t = 0; t < L::QTW; ++t) {
      const int qt = wave + 4 * t;
      if (qt >= L::NQT) break;
      f4 acc[NP];
#pragma unroll
      for (int b = 0; b < NP; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
      const int qa = qt * 16 + col;
#pragma unroll 7
      for (int kt = 0; kt < L::KTV; ++kt) {
        const int j = kt * 4 + kr;
        const bool ok = j < Np;
        const float a = (ok && qa < Nq) ? Vq[qa + j * Nq] : 0.f;
#pragma unroll
        for (int b = 0; b < NP; ++b) {
          const float bv = ok ? sR[col * RS + b * Np + j] : 0.f;
          acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv, acc[b], 0, 0, 0);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        S[q * SX + col] = acc[0][r] * c2r[t][r];
#pragma unroll
        for (int b = 1; b < NP; ++b) S[q * SX + b * 16 + col] = q < Nq ? acc[b][r] : 0.f;
      }
    }
  }
  __syncthreads();

  // ---- 8: rhs = Pq . S on MFMA, written back into the rhs tile
  {
    const int col = lane & 15, kr = lane >> 4;
    for (int mt = wave; mt < L::NPT; mt += 4) {
      f4 acc[NP];
#pragma unroll
      for (int b = 0; b < NP; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
      const int na = mt * 16 + col;
#pragma unroll 8
      for (int kt = 0; kt < L::KTP; ++kt) {
        const int q = kt * 4 + kr;
        const float a = (na < Np && q < Nq) ? Pq[na + q * Np] : 0.f;
#pragma unroll
        for (int b = 0; b < NP; ++b)
          acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, S[q * SX + b * 16 + col], acc[b], 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int b = 0; b < NP; ++b) sR[col * RS + b * Np + n] = acc[b][r];
        }
      }
    }
  }
  __syncthreads();

  // ---- 9: LSRK stage, flat and coalesced
  const float ft = ftime_dev ? *ftime_dev : ftime;
  const size_t base = (size_t)k0 * Nf * Np;
  for (int i = tid; i < nk * Nf * Np; i += 256) {
    const int e = i / (Nf * Np), rem = i - e * Nf * Np;
    float v = sR[e * RS + rem];
    const int slot = sSlot[e];
    if (slot >= 0) {
      const int f = rem / Np, n = rem - f * Np;
      v += src.w[f] * ft * src.coef[(size_t)slot * Np + n];
    }
    const float x = fa * res[base + i] + fdt * v;
    res[base + i] = x;
    Qout[base + i] = sQ[i] + fb * x;
  }
}

template <int N, bool P>
static hipError_t fused_mfma_n(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                               float fdt, const float *ftd, float ft, const SrcArgs &s) {
  using L = FusedLayout<N, P>;
  const size_t sh = sizeof(float) * L::floats;
  const int grid = (c->K + L::TB - 1) / L::TB;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused_mfma<N, P>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_fused_mfma<N, P>), dim3(grid), dim3(256), sh, c->stream, c->K, c->vgeo,
                     c->fgeo, c->vmapP, c->Fmask, c->Dids, c->Dvals, c->EEL_ids, c->EEL_vals,
                     c->L0_ids, c->L0_vals, c->Vq, c->Pq, c->c2q, fa, fb, fdt, s, ftd, ft, Qin,
                     c->res, Qout);
  return hipGetLastError();
}

template <int N>
static hipError_t fused_mfma_sel(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                                 float fdt, const float *ftd, float ft, int projU,
                                 const SrcArgs &s) {
  return projU ? fused_mfma_n<N, true>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s)
               : fused_mfma_n<N, false>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s);
}

hipError_t launch_fused_mfma(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                             float fdt, const float *ftd, float ft, int projU) {
  SrcArgs s;
  s.nsrc = c->nsrc;
  s.first = c->src_first;
  s.last = c->src_last;
  s.elems = c->src_elems;
  s.coef = c->src_coef;
  for (int f = 0; f < 9; ++f) s.w[f] = c->src_w[f];
  switch (c->N) {
    case 1: return fused_mfma_sel<1>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s);
    case 2: return fused_mfma_sel<2>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s);
    case 3: return fused_mfma_sel<3>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s);
    case 4: return fused_mfma_sel<4>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s);
    case 5: return fused_mfma_sel<5>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s);
    case 6: return fused_mfma_sel<6>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
