// Discrete adjoint of the acoustic fused LSRK45 stage (gfx950).
//
// Forward stage (W = Pq diag(c2) Vq on p, identity on u, v, w; A = volume + surface):
//     r = W A Q + s(t);  res' = a res + dt r;  Q' = Q + b res'
// Reverse stage, given the adjoints lQ (of Q') and lres (of res'):
//     g = lres + b lQ;  lres <- a g;  z = W^T g;  grad_c2 += dt (Pq^T g_p) * y;
//     lQ <- lQ + dt A^T z,          y = Vq (A Q)_p at the stage input ((A Q)_p saved by the forward)
// Kernels:
//   k_adj_wadg   g, lres update, t = Pq^T g_p, gradient, z_p = Vq^T (c2 t), z_u = g_u -> Z
//   k_adj_lift   h = 1/2 Fscale (L0^T EEL^T z_p - n . L0^T EEL^T z_u) per face node -> H
//   k_adj_apply  A^T z: volume transpose (D_m^T rows) + own and neighbour flux transposes
//                gathered through Fmask^-1 and mapP; lQ += dt A^T z
//   k_adj_inject receiver residuals into lQ_p (B_r^T (p_r - d_r))
#include <type_traits>

#include "bbw_fused_tile.h"

namespace bbw {

template <int N>
__global__ __launch_bounds__(256) void k_adj_wadg(int K, int KB, const float *__restrict__ Vq,
                                                  const float *__restrict__ Pq,
                                                  const float *__restrict__ c2q,
                                                  const float *__restrict__ Y, float fa, float fb,
                                                  float fdt, const float *__restrict__ lQ,
                                                  float *__restrict__ lres,
                                                  float *__restrict__ Z,
                                                  float *__restrict__ grad) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq, Nf = 4;
  extern __shared__ float lds[];
  float *sg = lds;                 // [KB][4][Np]
  float *st = lds + KB * Nf * Np;  // [KB][Nq]
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  const size_t base = (size_t)k0 * Nf * Np;
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) {
    const float g = lres[base + i] + fb * lQ[base + i];
    lres[base + i] = fa * g;
    sg[i] = g;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Nq; w += blockDim.x) {
    const int e = w / Nq, q = w - e * Nq;
    const float *g = sg + e * Nf * Np;
    const float *u = Y + (size_t)(k0 + e) * Np;
    float t = 0.f, y = 0.f;
    for (int n = 0; n < Np; ++n) {
      t += Pq[n + q * Np] * g[n];
      y += Vq[q + n * Nq] * u[n];
    }
    const size_t id = (size_t)(k0 + e) * Nq + q;
    grad[id] += fdt * t * y;
    st[e * Nq + q] = c2q[id] * t;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *s = st + e * Nq;
    float z = 0.f;
    for (int q = 0; q < Nq; ++q) z += Vq[q + n * Nq] * s[q];
    const float *g = sg + e * Nf * Np;
    float *o = Z + base + (size_t)e * Nf * Np;
    o[n] = z;
    o[Np + n] = g[Np + n];
    o[2 * Np + n] = g[2 * Np + n];
    o[3 * Np + n] = g[3 * Np + n];
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_adj_lift(int K, int KB, const float *__restrict__ fgeo,
                                                  const int *__restrict__ EELT_ids,
                                                  const float *__restrict__ EELT_vals, int WE,
                                                  const int *__restrict__ L0T_ids,
                                                  const float *__restrict__ L0T_vals, int WL,
                                                  const float *__restrict__ Z,
                                                  float *__restrict__ H) {
  constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, Nf = 4;
  extern __shared__ float lds[];
  float *sZ = lds;                    // [KB][4][Np]
  float *sT = sZ + KB * Nf * Np;      // [KB][4][NfpNf]  EEL^T z
  float *sF = sT + KB * Nf * NfpNf;   // [KB][16]
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) sZ[i] = Z[(size_t)k0 * Nf * Np + i];
  for (int i = threadIdx.x; i < nk * 16; i += blockDim.x) sF[i] = fgeo[(size_t)k0 * 16 + i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
    const int e = w / NfpNf, r = w - e * NfpNf;
    const float *z = sZ + e * Nf * Np;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    for (int j = 0; j < WE; ++j) {
      const int c = EELT_ids[r + j * NfpNf];
      const float v = EELT_vals[r + j * NfpNf];
      a0 += v * z[c];
      a1 += v * z[Np + c];
      a2 += v * z[2 * Np + c];
      a3 += v * z[3 * Np + c];
    }
    float *t = sT + e * Nf * NfpNf;
    t[r] = a0;
    t[NfpNf + r] = a1;
    t[2 * NfpNf + r] = a2;
    t[3 * NfpNf + r] = a3;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
    const int e = w / NfpNf, r = w - e * NfpNf, f = r / Nfp, n = r - f * Nfp;
    const float *t = sT + e * Nf * NfpNf + f * Nfp;
    float gp = 0.f, gu = 0.f, gv = 0.f, gw = 0.f;
    for (int j = 0; j < WL; ++j) {
      const int c = L0T_ids[n + j * Nfp];
      const float v = L0T_vals[n + j * Nfp];
      gp += v * t[c];
      gu += v * t[NfpNf + c];
      gv += v * t[2 * NfpNf + c];
      gw += v * t[3 * NfpNf + c];
    }
    const float *F = sF + 16 * e + 4 * f;
    const float gU = F[1] * gu + F[2] * gv + F[3] * gw;
    H[(size_t)(k0 + e) * NfpNf + r] = .5f * F[0] * (gp - gU);
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_adj_apply(int K, int KB, const float *__restrict__ vgeo,
                                                   const float *__restrict__ fgeo,
                                                   const int4 *__restrict__ DTids,
                                                   const float4 *__restrict__ DTvals,
                                                   const int *__restrict__ FmaskInv,
                                                   const int *__restrict__ mapP,
                                                   const float *__restrict__ Z,
                                                   const float *__restrict__ H, float fdt,
                                                   float *__restrict__ lQ) {
  constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, Nf = 4;
  extern __shared__ float lds[];
  float *sZ = lds;                // [KB][4][Np]: z_p, then Zr, Zs, Zt
  float *sG = sZ + KB * Nf * Np;  // [KB][9]
  float *sF = sG + KB * 9;        // [KB][16]
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) sZ[i] = Z[(size_t)k0 * Nf * Np + i];
  for (int i = threadIdx.x; i < nk * 9; i += blockDim.x) sG[i] = vgeo[(size_t)k0 * 9 + i];
  for (int i = threadIdx.x; i < nk * 16; i += blockDim.x) sF[i] = fgeo[(size_t)k0 * 16 + i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    float *s = sZ + e * Nf * Np;
    const float *G = sG + 9 * e;
    const float u = s[Np + n], v = s[2 * Np + n], ww = s[3 * Np + n];
    s[Np + n] = G[0] * u + G[1] * v + G[2] * ww;
    s[2 * Np + n] = G[3] * u + G[4] * v + G[5] * ww;
    s[3 * Np + n] = G[6] * u + G[7] * v + G[8] * ww;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np, k = k0 + e;
    const float *zp = sZ + e * Nf * Np, *Zr = zp + Np, *Zs = zp + 2 * Np, *Zt = zp + 3 * Np;
    const float *G = sG + 9 * e;
    float dp[4], dz[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int4 id = DTids[m * Np + n];
      const float4 vv = DTvals[m * Np + n];
      dp[m] = vv.x * zp[id.x] + vv.y * zp[id.y] + vv.z * zp[id.z] + vv.w * zp[id.w];
      if (m == 0)
        dz[0] = vv.x * (Zr[id.x] + Zs[id.x] + Zt[id.x]) + vv.y * (Zr[id.y] + Zs[id.y] + Zt[id.y]) +
                vv.z * (Zr[id.z] + Zs[id.z] + Zt[id.z]) + vv.w * (Zr[id.w] + Zs[id.w] + Zt[id.w]);
      else {
        const float *Zm = m == 1 ? Zr : m == 2 ? Zs : Zt;
        dz[m] = vv.x * Zm[id.x] + vv.y * Zm[id.y] + vv.z * Zm[id.z] + vv.w * Zm[id.w];
      }
    }
    // volume transpose
    float op = -.5f * (dz[1] + dz[2] + dz[3] - dz[0]);
    const float wr = .5f * (dp[1] - dp[0]), ws = .5f * (dp[2] - dp[0]), wt = .5f * (dp[3] - dp[0]);
    float ou = -(G[0] * wr + G[3] * ws + G[6] * wt);
    float ov = -(G[1] * wr + G[4] * ws + G[7] * wt);
    float ow = -(G[2] * wr + G[5] * ws + G[8] * wt);
    // flux transpose: own face-node slots of this node and their neighbours
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int r = FmaskInv[3 * n + a];
      if (r < 0) continue;
      const size_t own = (size_t)k * NfpNf + r;
      const float h = H[own];
      const int mp = mapP[own];
      const float *F = sF + 16 * e + 4 * (r / Nfp);
      if ((size_t)mp == own) {
        op -= 2.f * h;
      } else {
        const float hn = H[mp];
        op += hn - h;
        const float hs = h + hn;
        ou += hs * F[1];
        ov += hs * F[2];
        ow += hs * F[3];
      }
    }
    float *o = lQ + (size_t)k * Nf * Np + n;
    o[0] += fdt * op;
    o[Np] += fdt * ou;
    o[2 * Np] += fdt * ov;
    o[3 * Np] += fdt * ow;
  }
}

__global__ void k_adj_inject(int nrec, int Np, int NfNp, const int *__restrict__ elems,
                             const float *__restrict__ basis, const float *__restrict__ resid,
                             float *__restrict__ lQ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrec * Np) return;
  const int r = i / Np, j = i - r * Np;
  atomicAdd(lQ + (size_t)elems[r] * NfNp + j, basis[i] * resid[r]);
}

template <int N>
static hipError_t adj_stage_n(bbw_ctx *c, const float *Y, float fa, float fb, float fdt) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq, NfpNf = Dims<N>::NfpNf;
  const int KBw = 4, KBl = 4, KBa = 4;
  hipLaunchKernelGGL(k_adj_wadg<N>, dim3((c->K + KBw - 1) / KBw), dim3(256),
                     sizeof(float) * KBw * (4 * Np + Nq), c->stream, c->K, KBw, c->Vq, c->Pq,
                     c->c2q, Y, fa, fb, fdt, c->lQ, c->lres, c->Zadj, c->grad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_adj_lift<N>, dim3((c->K + KBl - 1) / KBl), dim3(256),
                     sizeof(float) * KBl * (4 * Np + 4 * NfpNf + 16), c->stream, c->K, KBl, c->fgeo,
                     c->EELT_ids, c->EELT_vals, c->WE, c->L0T_ids, c->L0T_vals, c->WL, c->Zadj,
                     c->Hadj);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_adj_apply<N>, dim3((c->K + KBa - 1) / KBa), dim3(256),
                     sizeof(float) * KBa * (4 * Np + 25), c->stream, c->K, KBa, c->vgeo, c->fgeo,
                     c->DTids, c->DTvals, c->FmaskInv, c->mapP, c->Zadj, c->Hadj, fdt, c->lQ);
  return hipGetLastError();
}

hipError_t launch_adj_stage(bbw_ctx *c, const float *Y, float fa, float fb, float fdt) {
  switch (c->N) {
    case 1: return adj_stage_n<1>(c, Y, fa, fb, fdt);
    case 2: return adj_stage_n<2>(c, Y, fa, fb, fdt);
    case 3: return adj_stage_n<3>(c, Y, fa, fb, fdt);
    case 4: return adj_stage_n<4>(c, Y, fa, fb, fdt);
    case 5: return adj_stage_n<5>(c, Y, fa, fb, fdt);
    case 6: return adj_stage_n<6>(c, Y, fa, fb, fdt);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_adj_inject(bbw_ctx *c, const float *resid_row) {
  const int n = c->nrec * c->Np;
  hipLaunchKernelGGL(k_adj_inject, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->nrec, c->Np,
                     c->Nf * c->Np, c->rec_elems, c->rec_basis, resid_row,
                     c->lQ + (size_t)c->rec_field * c->Np);
  return hipGetLastError();
}

}  // namespace bbw

// ---------------------------------------------------------------------------------------------
// Tiled adjoint stage: two kernels per reverse stage, 16-element tiles, W^T on FP32 MFMA.
namespace bbw {

typedef float f4a __attribute__((ext_vector_type(4)));

template <int N>
struct AdjLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf;
  static constexpr int Nq = Dims<N>::Nq, TB = 16;
  static constexpr int NQT = (Nq + 15) / 16, NQP = 16 * NQT, NPT = (Np + 15) / 16, NPP = 16 * NPT;
  static constexpr int KTV = (Np + 3) / 4, KTP = NQP / 4, QTW = (NQT + 3) / 4;
  static constexpr int KBV = (KTV + 3) / 4, KBP = (KTP + 3) / 4;
  static constexpr int FMT = (Nfp + 15) / 16;      // face-node M tiles per face
  static constexpr int NPR = 4 * FMT, PRW = (NPR + 3) / 4;   // (face, M tile) pairs; per wave
  // g / z rows [TB][RS]: 4 fields, stride 2 mod 32 for the MFMA operand reads
  static constexpr int RS0 = 4 * Np, RS = RS0 + ((34 - RS0 % 32) % 32);
  static constexpr int NG3 = 256 / NfpNf, EPT3 = (TB + NG3 - 1) / NG3;
  static constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS, EPTV = (TB + NG - 1) / NG;
  // degree N-1 nodes of the factored transposed derivative (k_adj2 step 1)
  static constexpr int Np1 = N * (N + 1) * (N + 2) / 6;
  static constexpr int GS1 = Np1 <= 32 ? 32 : 64, NG1 = 256 / GS1, EPT1 = TB / NG1;
  static_assert(Np1 <= 64 && NG1 * EPT1 == TB && 5 * Np1 <= 4 * Np, "step 1 covers the tile and fits sZ");
  static constexpr bool PF = 4 * Np <= 256;
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  // sg [TB][RS] | S [NQP][16] | sF [TB][16]
  static constexpr int U = NQP * 16;
  // (A Q)_p rows [TB][RU]: K padded to KU, stride 2 mod 32
  static constexpr int KU = 4 * KTV, RU = KU + ((34 - KU % 32) % 32);
  // c2 and gradient tiles [TB * Nq] each (contiguous copies of the global rows); the H tile
  // [TB][NfpNf] reuses them
  // (padded to whole 256-word rounds, so every thread stores every word it loads: a conditional
  // store lets the compiler sink its load, serialising the rounds)
  static constexpr int CQ = ((TB * Nq + 255) / 256) * 256;
  static constexpr int CG = cmax(2 * CQ, TB * NfpNf);
  static constexpr int NIU = (TB * Np + 255) / 256, NIC = CQ / 256;
  // ... | one scratch word for the (A Q)_p words past the tile
  static constexpr size_t floats1 = (size_t)TB * RS + U + TB * 16 + (size_t)TB * RU + CG + 1;
  // sZ4 float4 [TB][Np] | sH [TB][NfpNf] padded to whole 256-word rounds | sG [TB][9] in a
  // 256-word slot | sF [TB][16]
  static constexpr int HIT = (TB * NfpNf + 255) / 256;
  static constexpr size_t floats2 = (size_t)TB * 4 * Np + (size_t)HIT * 256 + 256 + TB * 16;
};

// One reverse stage, element-local part, on 16-element tiles:
//   g = lres + b lQ; lres <- a g; t = Pq^T g_p (MFMA); grad += dt t y; z_p = Vq^T (c2 t) (MFMA);
//   z_u = g_u -> Z;  h_f = 1/2 Fscale_f LIFT_f^T (z_p - n_f . z_u) per face (MFMA) -> H
// (LIFT_f^T = L0^T EEL^T restricted to face f, applied densely on the matrix cores).
template <int N>
__global__ __launch_bounds__(256, N <= 5 ? 4 : 2) void k_adj1(int K, FusedOps o, const float *__restrict__ c2q,
                                              const float *__restrict__ Y,
                                              const float *__restrict__ fgeo, float fa, float fb,
                                              float fdt, int lres_io, const float *__restrict__ lQ,
                                              float *__restrict__ lres, float *__restrict__ Z,
                                              float *__restrict__ H, float *__restrict__ grad,
                                              const float *__restrict__ PqTBlk,
                                              const float *__restrict__ VqTBlk,
                                              const float *__restrict__ LiftTBlk, int xcd_chunk) {
  using L = AdjLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, Nq = L::Nq, TB = L::TB, RS = L::RS;
  constexpr int NpNf = 4 * Np, KTV = L::KTV, KTP = L::KTP, KBV = L::KBV, KBP = L::KBP;
  constexpr int FMT = L::FMT;
  constexpr unsigned ROW = NpNf * 4u;
  extern __shared__ float lds[];
  float *sg = lds;                        // [TB][RS]
  float *S = sg + TB * RS;                // [NQP][16]   (phases 2-3)
  float *sF = S + L::U;                   // [TB][16]
  float *sU = sF + TB * 16;               // [TB][RU] (A Q)_p of the forward stage
  float *sC = sU + TB * L::RU;            // [TB * Nq] c2        (phases 1-2)
  float *sGr = sC + L::CQ;                // [TB * Nq] gradient  (phases 1-3)
  float *sScr = sC + L::CG;               // scratch word
  float *sHt = sC;                        // [TB][NfpNf] H       (phases 4-5)
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = tile_index(xcd_chunk) * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * NpNf;
  const unsigned tbytes = (unsigned)nk * ROW, qbytes = (unsigned)nk * Nq * 4u;
  const brsrc rQl = mkrsrc(lQ + base, tbytes), rRl = mkrsrc(lres + base, tbytes);
  const brsrc rZ = mkrsrc(Z + base, tbytes);
  const brsrc rH = mkrsrc(H + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);
  const brsrc rC = mkrsrc(c2q + (size_t)k0 * Nq, qbytes);
  const brsrc rY = mkrsrc(Y + (size_t)k0 * Np, (unsigned)nk * Np * 4u);
  const brsrc rVq = mkrsrc(o.VqBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rGd = mkrsrc(grad + (size_t)k0 * Nq, qbytes);
  const brsrc rPT = mkrsrc(PqTBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rVT = mkrsrc(VqTBlk, (unsigned)L::NPT * KBP * 1024u);
  const brsrc rLT = mkrsrc(LiftTBlk, 4u * FMT * KBV * 1024u);

  // ---- 1: g = lres + b lQ -> LDS; lres <- a g; fgeo; c2, y, grad in the MFMA output layout.
  // lres_io bit 0: read lres (else it is 0); bit 1: store a g (else it is dead).  Every load of
  // the phase is issued before the first store; the stores are unconditional (lanes past a row
  // read and write through an out-of-range offset, which returns 0 and drops the write, and
  // store to LDS at the scratch word).
  const float fval = bld(mkrsrc(fgeo + (size_t)k0 * 16, (unsigned)nk * 64u), tid * 4u);
  const bool st_lres = (lres_io & 2) != 0;
  const unsigned gof = tid < NpNf ? tid * 4u : 0x80000000u;
  // the raw rows stay in registers until the stores: g formed right after its loads would make
  // the later loads of the phase wait for them
  float lrv[L::PF ? TB : 1], lqv[L::PF ? TB : 1];
  if constexpr (L::PF) {
    const unsigned rof = (lres_io & 1) ? gof : 0x80000000u;
#pragma unroll
    for (int e = 0; e < TB; ++e) {
      lrv[e] = bld(rRl, rof + e * ROW);
      lqv[e] = bld(rQl, gof + e * ROW);
    }
  } else {
    for (int rr = tid; rr < NpNf; rr += 256)
      for (int e = 0; e < TB; ++e) {
        const float lr = (lres_io & 1) ? bld(rRl, rr * 4u + e * ROW) : 0.f;
        const float g = lr + fb * bld(rQl, rr * 4u + e * ROW);
        sg[e * RS + rr] = g;
        if (st_lres) bst(rRl, fa * g, rr * 4u + e * ROW);
      }
  }
  // (A Q)_p rows, c2 and gradient rows of the tile (the last two contiguous in memory and copied
  // as such), then the operand blocks of phase 2
  float yv[L::NIU], cv[L::NIC], gvv[L::NIC];
#pragma unroll
  for (int j = 0; j < L::NIU; ++j) yv[j] = bld(rY, (unsigned)(tid + 256 * j) * 4u);
#pragma unroll
  for (int j = 0; j < L::NIC; ++j) {
    cv[j] = bld(rC, (unsigned)(tid + 256 * j) * 4u);
    gvv[j] = bld(rGd, (unsigned)(tid + 256 * j) * 4u);
  }
  f4 aT[L::QTW][KBV], aY[L::QTW][KBV];
  // (under the wave test: loading the last tile unconditionally instead, as the MFMA updates
  // do, made this kernel slower: 214 against 205 us per stage on 196608 tets, N = 5)
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t;
    if (qt < L::NQT) {
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb) {
        aT[t][kb] = bld_blk<KTV>(rPT, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
        aY[t][kb] = bld_blk<KTV>(rVq, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
      }
    }
  }
  sF[tid] = fval;
  if constexpr (L::PF) {
    float gv[TB];
#pragma unroll
    for (int e = 0; e < TB; ++e) gv[e] = lrv[e] + fb * lqv[e];
    float *dst = tid < NpNf ? sg + tid : sScr;
    const int step = tid < NpNf ? RS : 0;
#pragma unroll
    for (int e = 0; e < TB; ++e) dst[e * step] = gv[e];
    if (st_lres) {
#pragma unroll
      for (int e = 0; e < TB; ++e) bst(rRl, fa * gv[e], gof + e * ROW);
    }
  }
#pragma unroll
  for (int j = 0; j < L::NIU; ++j) {
    const int i = tid + 256 * j, e = i / Np;
    float *dst = i < TB * Np ? sU + e * L::RU + (i - e * Np) : sScr;
    *dst = yv[j];
  }
  if (tid < TB * (L::KU - Np)) sU[(tid / (L::KU - Np)) * L::RU + Np + tid % (L::KU - Np)] = 0.f;
#pragma unroll
  for (int j = 0; j < L::NIC; ++j) {
    sC[tid + 256 * j] = cv[j];
    sGr[tid + 256 * j] = gvv[j];
  }
  __syncthreads();

  // ---- 2: t = Pq^T g_p and y = Vq (A Q)_p (M = cubature points, N = elements, K = nodes);
  // grad += dt t y; S = c2 t
  f4 aV[KBP];
  {
    const int mt0 = wave < L::NPT ? wave : L::NPT - 1;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aV[kb] = bld_blk<KTP>(rVT, lane * 16u, (unsigned)(mt0 * KBP + kb) * 1024u, kb);
  }
  {
    const float *Bp = sg + col * RS + kr;
    const float *Bu = sU + col * L::RU + kr;
#pragma unroll
    for (int t = 0; t < L::QTW; ++t) {
      const int qt = wave + 4 * t;
      if (qt >= L::NQT) break;
      f4 acc = {0.f, 0.f, 0.f, 0.f}, ay = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb)
          if (4 * kb + sb < KTV) {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(aT[t][kb][sb], Bp[(4 * kb + sb) * 4], acc, 0, 0, 0);
            ay = __builtin_amdgcn_mfma_f32_16x16x4f32(aY[t][kb][sb], Bu[(4 * kb + sb) * 4], ay, 0, 0, 0);
          }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        float sv = 0.f;
        if (q < Nq) {
          sGr[col * Nq + q] += fdt * acc[r] * ay[r];
          sv = sC[col * Nq + q] * acc[r];
        }
        S[q * 16 + col] = sv;
      }
    }
  }
  __syncthreads();

  // ---- 3: z_p = Vq^T S (M = nodes, K = cubature points), over g_p; gradient rows out
  f4 aL[L::PRW][KBV];
#pragma unroll
  for (int j = 0; j < L::PRW; ++j) {
    // unconditional, clamped to the last row tile (a wave past it loads that one and does not
    // use it): at N = 5 every wave has both tiles, and without the test the waits are exact
    const int pr = wave + 4 * j < L::NPR ? wave + 4 * j : L::NPR - 1;
#pragma unroll
    for (int kb = 0; kb < KBV; ++kb) aL[j][kb] = bld_blk<KTV>(rLT, lane * 16u, (unsigned)(pr * KBV + kb) * 1024u, kb);
  }
  for (int i = tid; i < TB * Nq; i += 256) bst(rGd, sGr[i], i * 4u);
  for (int mt = wave; mt < L::NPT; mt += 4) {
    if (mt != wave) {
#pragma unroll
      for (int kb = 0; kb < KBP; ++kb) aV[kb] = bld_blk<KTP>(rVT, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const float *Bp = S + kr * 16 + col;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb)
#pragma unroll
      for (int sb = 0; sb < 4; ++sb)
        if (4 * kb + sb < KTP)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(aV[kb][sb], Bp[(4 * kb + sb) * 64], acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = mt * 16 + 4 * kr + r;
      if (n < Np) sg[col * RS + n] = acc[r];
    }
  }
  __syncthreads();

  // ---- 4: Z out; h_f = 1/2 Fscale_f LIFT_f^T (z_p - n_f . z_u) per face on MFMA (M = face
  // nodes, N = elements, K = nodes; the B operand is formed from z in LDS) -> H tile in LDS
  if constexpr (L::PF) {
    if (tid < NpNf) {
#pragma unroll
      for (int e = 0; e < TB; ++e) bst(rZ, sg[e * RS + tid], tid * 4u + e * ROW);
    }
  } else {
    for (int rr = tid; rr < NpNf; rr += 256)
      for (int e = 0; e < TB; ++e) bst(rZ, sg[e * RS + rr], rr * 4u + e * ROW);
  }
  {
    const float *zc = sg + col * RS + kr;
    const float *Fe = sF + 16 * col;
#pragma unroll
    for (int j = 0; j < L::PRW; ++j) {
      const int pr = wave + 4 * j;
      if (pr >= L::NPR) break;
      const int f = pr / FMT, mt = pr - f * FMT;
      const float nx = Fe[4 * f + 1], ny = Fe[4 * f + 2], nz = Fe[4 * f + 3];
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < KTV) {
            const int n = 4 * kt;
            float w = zc[n] - (nx * zc[Np + n] + ny * zc[2 * Np + n] + nz * zc[3 * Np + n]);
            if (n + 3 >= Np && n + kr >= Np) w = 0.f;   // K padding past the last node
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(aL[j][kb][sb], w, acc, 0, 0, 0);
          }
        }
      const float hs = .5f * Fe[4 * f];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = mt * 16 + 4 * kr + r;
        if (m < Nfp) sHt[col * NfpNf + f * Nfp + m] = hs * acc[r];
      }
    }
  }
  __syncthreads();
  // ---- 5: H rows out (contiguous)
  for (int i = tid; i < TB * NfpNf; i += 256) bst(rH, sHt[i], i * 4u);
}

// lQ += dt A^T z: volume transpose from (z_p, Zr, Zs, Zt) float4 per node, flux transpose from
// own H and the neighbour's H through mapP.  Every global read (Z rows, lQ rows, mapP entries,
// neighbour H) is issued before the first barrier.
template <int N>
__global__ __launch_bounds__(256) void k_adj2(int K, const float *__restrict__ vgeo,
                                              const float *__restrict__ fgeo,
                                              const int4 *__restrict__ Sids,
                                              const float4 *__restrict__ Swts,
                                              const int4 *__restrict__ Eids,
                                              const float4 *__restrict__ Evals,
                                              const int *__restrict__ FmaskInv,
                                              const int *__restrict__ mapP,
                                              const float *__restrict__ Z,
                                              const float *__restrict__ H, float fdt,
                                              float *__restrict__ lQ, int xcd_chunk) {
  using L = AdjLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, TB = L::TB, NpNf = 4 * Np;
  constexpr unsigned ROW = NpNf * 4u;
  extern __shared__ float lds[];
  float4 *sZ = (float4 *)lds;              // [TB][Np]
  float *sH = (float *)(sZ + TB * Np);     // [TB][NfpNf], HIT x 256 words
  float *sG = sH + L::HIT * 256;           // [TB][9], 256 words
  float *sF = sG + 256;                    // [TB][16]
  const int tid = threadIdx.x;
  const int k0 = tile_index(xcd_chunk) * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * NpNf;
  const brsrc rZ = mkrsrc(Z + base, (unsigned)nk * ROW), rL = mkrsrc(lQ + base, (unsigned)nk * ROW);
  const brsrc rHt = mkrsrc(H + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);
  const brsrc rHa = mkrsrc(H, (unsigned)K * NfpNf * 4u);
  const brsrc rM = mkrsrc(mapP + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);
  // Loads in the order of first use, fixed by scheduling barriers (the memory counter is
  // in order, so waiting for a load also waits for every load issued before it; left to
  // itself the compiler issued the Z rows, which the first phase after the barrier reads,
  // last): the node's face slots (the mapP gather depends on them), the Z rows, the volume
  // tables, the H tile and geometry (stored to LDS before the barrier), the mapP entries, the
  // lQ rows (read at the end).  The tile goes to LDS with unconditional stores into whole
  // 256-word rounds (words past the tile hold zeros, never read).
  const int gv = tid / L::GS, nv = tid - gv * L::GS;
  const bool on = gv < L::NG && nv < Np;
  int slot[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) slot[a] = on ? FmaskInv[3 * nv + a] : -1;
  float zr[L::EPTV][4], lq[L::EPTV][4], hn[L::EPTV][3];
  int mp[L::EPTV][3];
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i)
#pragma unroll
    for (int f = 0; f < 4; ++f) zr[i][f] = bld(rZ, (unsigned)((gv + L::NG * i) * NpNf + f * Np + nv) * 4u);
  __builtin_amdgcn_sched_barrier(0);
  // the thread's rows of the shift and elevation tables (lanes past the nodes read row 0 and
  // do not use it)
  const int g1 = tid / L::GS1, n1 = tid - g1 * L::GS1;
  const int4 S = Sids[n1 < L::Np1 ? n1 : 0];
  const float4 wt = Swts[n1 < L::Np1 ? n1 : 0];
  const int4 E = Eids[nv < Np ? nv : 0];
  const float4 ev = Evals[nv < Np ? nv : 0];
  __builtin_amdgcn_sched_barrier(0);
  float hreg[L::HIT];
#pragma unroll
  for (int r = 0; r < L::HIT; ++r) hreg[r] = bld(rHt, (unsigned)(tid + 256 * r) * 4u);
  const float gval = bld(mkrsrc(vgeo + (size_t)k0 * 9, (unsigned)nk * 36u), tid * 4u);
  const float fval = bld(mkrsrc(fgeo + (size_t)k0 * 16, (unsigned)nk * 64u), tid * 4u);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a)
      mp[i][a] = bldi(rM, (unsigned)((gv + L::NG * i) * NfpNf + (slot[a] < 0 ? 0 : slot[a])) * 4u);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i)
#pragma unroll
    for (int f = 0; f < 4; ++f) lq[i][f] = bld(rL, (unsigned)((gv + L::NG * i) * NpNf + f * Np + nv) * 4u);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int r = 0; r < L::HIT; ++r) sH[tid + 256 * r] = hreg[r];
  sG[tid] = gval;
  sF[tid] = fval;
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) hn[i][a] = bld(rHa, (unsigned)mp[i][a] * 4u);
  __syncthreads();
  if (on) {
#pragma unroll
    for (int i = 0; i < L::EPTV; ++i) {
      const int e = gv + L::NG * i;
      const float *G = sG + 9 * e;
      const float u = zr[i][1], v = zr[i][2], w = zr[i][3];
      sZ[e * Np + nv] = make_float4(zr[i][0], G[0] * u + G[1] * v + G[2] * w,
                                    G[3] * u + G[4] * v + G[5] * w, G[6] * u + G[7] * v + G[8] * w);
    }
  }
  __syncthreads();
  // Volume transpose in factored form.  D_m = E . Shift_m (times N), so
  // D_m^T v [c] = [c_m >= 1] T[c - e_m] with T = N E^T v, T[b] = sum_j (b_j + 1) v[b + e_j] at the
  // degree-(N-1) nodes b.  Step 1 forms T of p, Ur + Us + Ut, Ur, Us and Ut from the (p, U) tile;
  // step 2 reads it at c - e_m.
  float4 ta[L::EPT1];
  float tb[L::EPT1];
  if (n1 < L::Np1) {
#pragma unroll
    for (int i = 0; i < L::EPT1; ++i) {
      const float4 *V = sZ + (g1 + L::NG1 * i) * Np;
      const float4 q1 = V[S.x], q2 = V[S.y], q3 = V[S.z], q4 = V[S.w];
      const float ur = wt.x * q1.y + wt.y * q2.y + wt.z * q3.y + wt.w * q4.y;
      const float us = wt.x * q1.z + wt.y * q2.z + wt.z * q3.z + wt.w * q4.z;
      const float ut = wt.x * q1.w + wt.y * q2.w + wt.z * q3.w + wt.w * q4.w;
      ta[i] = make_float4(wt.x * q1.x + wt.y * q2.x + wt.z * q3.x + wt.w * q4.x, ur + us + ut, ur, us);
      tb[i] = ut;
    }
  }
  __syncthreads();
  float4 *sT = sZ;                                     // [TB][Np1] (p, Ur + Us + Ut, Ur, Us)
  float *sTw = (float *)(sT + TB * L::Np1);            // [TB][Np1] Ut
  if (n1 < L::Np1) {
#pragma unroll
    for (int i = 0; i < L::EPT1; ++i) {
      const int e = g1 + L::NG1 * i;
      sT[e * L::Np1 + n1] = ta[i];
      sTw[e * L::Np1 + n1] = tb[i];
    }
  }
  __syncthreads();
  if (!on) return;
  const float mx = ev.x > 0.f ? 1.f : 0.f, my = ev.y > 0.f ? 1.f : 0.f;
  const float mz = ev.z > 0.f ? 1.f : 0.f, mw = ev.w > 0.f ? 1.f : 0.f;
#pragma unroll
  for (int i = 0; i < L::EPTV; ++i) {
    const int e = gv + L::NG * i;
    const float4 *T = sT + e * L::Np1;
    const float4 t1 = T[E.x], t2 = T[E.y], t3 = T[E.z], t4 = T[E.w];
    const float t4w = sTw[e * L::Np1 + E.w];
    float dp[4], dz[4];
    dp[0] = mx * t1.x; dz[0] = mx * t1.y;
    dp[1] = my * t2.x; dz[1] = my * t2.z;
    dp[2] = mz * t3.x; dz[2] = mz * t3.w;
    dp[3] = mw * t4.x; dz[3] = mw * t4w;
    const float *G = sG + 9 * e;
    float op = -.5f * (dz[1] + dz[2] + dz[3] - dz[0]);
    const float wr = .5f * (dp[1] - dp[0]), ws = .5f * (dp[2] - dp[0]), wt = .5f * (dp[3] - dp[0]);
    float ou = -(G[0] * wr + G[3] * ws + G[6] * wt);
    float ov = -(G[1] * wr + G[4] * ws + G[7] * wt);
    float ow = -(G[2] * wr + G[5] * ws + G[8] * wt);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int r = slot[a];
      if (r < 0) continue;
      const int own = (k0 + e) * NfpNf + r;
      const float h = sH[e * NfpNf + r];
      const float *F = sF + 16 * e + 4 * (r / Nfp);
      if (mp[i][a] == own) {
        op -= 2.f * h;
      } else {
        op += hn[i][a] - h;
        const float hs = h + hn[i][a];
        ou += hs * F[1];
        ov += hs * F[2];
        ow += hs * F[3];
      }
    }
    const unsigned o = (unsigned)(e * NpNf + nv) * 4u;
    bst(rL, lq[i][0] + fdt * op, o);
    bst(rL, lq[i][1] + fdt * ou, o + Np * 4u);
    bst(rL, lq[i][2] + fdt * ov, o + 2 * Np * 4u);
    bst(rL, lq[i][3] + fdt * ow, o + 3 * Np * 4u);
  }
}

template <int N>
static hipError_t adj_tiled_n(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io) {
  using L = AdjLayout<N>;
  const int grid = (c->K + L::TB - 1) / L::TB;
  const size_t s1 = sizeof(float) * L::floats1, s2 = sizeof(float) * L::floats2;
  if (s1 > 65536)
    hipFuncSetAttribute((const void *)k_adj1<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s1);
  if (s2 > 65536)
    hipFuncSetAttribute((const void *)k_adj2<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s2);
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_adj1<N>, dim3(grid), dim3(256), s1, c->stream, c->K, fused_ops(c), c->c2q, Y,
                     c->fgeo, fa, fb, fdt, lres_io, c->lQ, c->lres, c->Zadj, c->Hadj, c->grad, c->PqTBlk,
                     c->VqTBlk, c->LiftTBlk, c->xcd_chunk);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_adj2<N>, dim3(grid), dim3(256), s2, c->stream, c->K, c->vgeo, c->fgeo,
                     c->Sids, c->Swts, c->Eids, c->Evals, c->FmaskInv, c->mapP, c->Zadj, c->Hadj, fdt,
                     c->lQ, c->xcd_chunk);
  return hipGetLastError();
}

hipError_t launch_adj_stage_tiled(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io) {
  switch (c->N) {
    case 1: return adj_tiled_n<1>(c, Y, fa, fb, fdt, lres_io);
    case 2: return adj_tiled_n<2>(c, Y, fa, fb, fdt, lres_io);
    case 3: return adj_tiled_n<3>(c, Y, fa, fb, fdt, lres_io);
    case 4: return adj_tiled_n<4>(c, Y, fa, fb, fdt, lres_io);
    case 5: return adj_tiled_n<5>(c, Y, fa, fb, fdt, lres_io);
    case 6: return adj_tiled_n<6>(c, Y, fa, fb, fdt, lres_io);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
