// Discrete adjoint of the fused elastic LSRK45 stage (gfx950), two kernels per reverse stage on
// 16-element tiles.  No counterpart in the reference; the forward stage is rk_volume_elas +
// rk_surface_elas + rk_update_elas (elas_cpp/okl/ElasKernelsWADG.okl:25-433) with the
// Bernstein-Bezier operators, fields v1 v2 v3 sxx syy szz syz sxz sxy:
//     r = W(m) A Q + s(t);  res' = a res + dt r;  Q' = Q + b res'
// with W(m) = Pq M(m) Vq per field group (1/rho on the velocities; the isotropic stiffness
// (lambda, mu) on the normal stresses; mu on the shear stresses; M symmetric at every point).
// Reverse stage, given the adjoints lQ (of Q') and lres (of res'):
//     g = lres + b lQ;  lres <- a g;  t = Pq^T g;  y = Vq (A Q) (A Q saved by the forward);
//     grad(1/rho) += dt sum_v t y;  grad(lambda) += dt (sum_n t)(sum_n y);
//     grad(mu) += dt (2 sum_n t y + sum_s t y)   (v velocities, n normal, s shear stresses);
//     z = Vq^T M(m) t;  lQ <- lQ + dt A^T z
// k_adje1: everything up to z on MFMA, then the face part of A^T: w = 1/2 Fscale LIFT_f^T z per
//   face node (MFMA) and the transpose of the penalty flux.  The flux on a face is
//   fl_v = S^T dsigma + tau_v S^T S dv and fl_sigma = S (dv + tau_s S^T dsigma), with S = S(n) the
//   3 -> 6 map a -> (a1 nx, a2 ny, a3 nz, a3 ny + a2 nz, a3 nx + a1 nz, a2 nx + a1 ny), so its
//   transpose is h_v = b + tau_v S^T S w_v, h_sigma = S c with b = S^T w_sigma, c = w_v + tau_s b:
//   six numbers (h_v, c) per face node -> H [K][4Nfp][6].
// k_adje2: volume transpose in factored form (D_m^T v [c] = [c_m >= 1] T[c - e_m], T = N E^T v at
//   the degree N-1 nodes) for the 9 fields, and the flux transpose: the own trace gets -h, the
//   neighbour's +h (through the face-node map; its S(n) is -S(n) of the own face); lQ += dt A^T z.
//
// v_mfma_f32_16x16x4f32 lane layout: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k,
// D[4*(l/16) + r][l%16] in register r of lane l.
#include "bbw_fused_tile.h"

namespace bbw {

template <int N>
struct AdjElasLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf;
  static constexpr int Nq = Dims<N>::Nq, TB = 16, RowF = 9 * Np;
  static constexpr int NQT = (Nq + 15) / 16, NQP = 16 * NQT, NPT = (Np + 15) / 16;
  static constexpr int KTV = (Np + 3) / 4, KTP = NQP / 4, QTW = (NQT + 3) / 4;
  static constexpr int KBV = (KTV + 3) / 4, KBP = (KTP + 3) / 4;
  static constexpr int FMT = (Nfp + 15) / 16, NPR = 4 * FMT;   // (face, node tile) pairs
  // g / z and (A Q) rows [TB][RS]: 9 fields plus the K padding the products read past the
  // ninth, stride 2 mod 32 (MFMA operand reads: 16 rows x 2 consecutive words)
  static constexpr int RS0 = RowF > 8 * Np + 4 * KTV ? RowF : 8 * Np + 4 * KTV;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  // k_adje1: sg [TB][RS] | sU [TB][RS] | S [3][NQP][16] | sF [TB][16] | one scratch word;
  // the H tile [TB][NfpNf][6] reuses sU and S
  static constexpr size_t floats1 = 2 * (size_t)TB * RS + 3 * NQP * 16 + TB * 16 + 1;
  static_assert(TB * NfpNf * 6 <= TB * RS + 3 * NQP * 16, "H tile fits sU + S");
  // k_adje2: sZ [TB][RowF] | T [TB][9][Np1] | face terms [TB][9][NfpNf] | vgeo, fgeo (256 each)
  static constexpr int Np1 = N * (N + 1) * (N + 2) / 6;
  static constexpr size_t floats2 = (size_t)TB * RowF + (size_t)TB * 9 * Np1 + (size_t)TB * 9 * NfpNf + 512;
};

struct AdjElasMedia {
  const float *rhoinv, *lambda, *mu;
  float tau_v, tau_s;
};

template <int N>
__global__ __launch_bounds__(256) void k_adje1(int K, FusedOps o, AdjElasMedia md,
                                               const float *__restrict__ Y,
                                               const float *__restrict__ fgeo, float fa, float fb,
                                               float fdt, int lres_io, const float *__restrict__ lQ,
                                               float *__restrict__ lres, float *__restrict__ Z,
                                               float *__restrict__ H, float *__restrict__ grad,
                                               const float *__restrict__ PqTBlk,
                                               const float *__restrict__ VqTBlk,
                                               const float *__restrict__ LiftTBlk) {
  using L = AdjElasLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, Nq = L::Nq, TB = L::TB, RS = L::RS;
  constexpr int RowF = L::RowF, NQP = L::NQP, KTV = L::KTV, KTP = L::KTP, KBV = L::KBV, KBP = L::KBP;
  constexpr unsigned ROW = RowF * 4u;
  extern __shared__ float lds[];
  float *sg = lds;                    // [TB][RS] g, then z
  float *sU = sg + TB * RS;           // [TB][RS] (A Q) of the forward stage
  float *S = sU + TB * RS;            // [3][NQP][16]
  float *sF = S + 3 * NQP * 16;       // [TB][16] fgeo
  float *sScr = sF + TB * 16;         // written by lanes past a row, never read
  float *sHt = sU;                    // [TB][NfpNf][6] (phase 4)
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * RowF, KNq = (size_t)K * Nq;
  const unsigned tbytes = (unsigned)nk * ROW, qbytes = (unsigned)nk * Nq * 4u;
  const brsrc rQl = mkrsrc(lQ + base, tbytes), rRl = mkrsrc(lres + base, tbytes);
  const brsrc rZ = mkrsrc(Z + base, tbytes), rY = mkrsrc(Y + base, tbytes);
  const brsrc rH = mkrsrc(H + (size_t)k0 * NfpNf * 6, (unsigned)nk * NfpNf * 24u);
  const brsrc rRho = mkrsrc(md.rhoinv + (size_t)k0 * Nq, qbytes);
  const brsrc rLam = mkrsrc(md.lambda + (size_t)k0 * Nq, qbytes);
  const brsrc rMu = mkrsrc(md.mu + (size_t)k0 * Nq, qbytes);
  const brsrc rG0 = mkrsrc(grad + (size_t)k0 * Nq, qbytes);
  const brsrc rG1 = mkrsrc(grad + KNq + (size_t)k0 * Nq, qbytes);
  const brsrc rG2 = mkrsrc(grad + 2 * KNq + (size_t)k0 * Nq, qbytes);
  const brsrc rVq = mkrsrc(o.VqBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rPT = mkrsrc(PqTBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rVT = mkrsrc(VqTBlk, (unsigned)L::NPT * KBP * 1024u);
  const brsrc rLT = mkrsrc(LiftTBlk, 4u * L::FMT * KBV * 1024u);

  // ---- 1: g = lres + b lQ -> sg, lres <- a g, (A Q) rows -> sU, fgeo; media and gradient at
  // the wave's cubature points in the MFMA output layout (lane: element col, points
  // qt*16 + 4 kr + r).  lres_io bit 0: read lres (else it is 0); bit 1: store a g (else dead).
  sF[tid] = bld(mkrsrc(fgeo + (size_t)k0 * 16, (unsigned)nk * 64u), tid * 4u);
  {
    constexpr int RR = (RowF + 255) / 256;
    const bool rd = (lres_io & 1) != 0, st = (lres_io & 2) != 0;
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      const int rr = tid + 256 * k;
      float lr[TB], lq[TB], yv[TB];
#pragma unroll
      for (int e = 0; e < TB; ++e) {
        const unsigned off = rr < RowF ? (unsigned)rr * 4u + e * ROW : 0x80000000u;
        lr[e] = bld(rRl, rd ? off : 0x80000000u);
        lq[e] = bld(rQl, off);
        yv[e] = bld(rY, off);
      }
      float *dg = rr < RowF ? sg + rr : sScr;
      float *du = rr < RowF ? sU + rr : sScr;
      const int step = rr < RowF ? RS : 0;
#pragma unroll
      for (int e = 0; e < TB; ++e) {
        const float g = lr[e] + fb * lq[e];
        dg[e * step] = g;
        du[e * step] = yv[e];
        if (st) bst(rRl, fa * g, rr < RowF ? (unsigned)rr * 4u + e * ROW : 0x80000000u);
      }
    }
  }
  constexpr int KPAD = L::RS0 - RowF;
  static_assert(TB * KPAD <= 256, "one padding word per thread");
  if (KPAD > 0 && tid < TB * KPAD) {
    const int i = (tid / (KPAD > 0 ? KPAD : 1)) * RS + RowF + tid % (KPAD > 0 ? KPAD : 1);
    sg[i] = 0.f;
    sU[i] = 0.f;
  }
  float rho[L::QTW][4], lam[L::QTW][4], mu[L::QTW][4], g0[L::QTW][4], g1[L::QTW][4], g2[L::QTW][4];
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t < L::NQT ? wave + 4 * t : L::NQT - 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const unsigned off = (unsigned)(col * Nq + qt * 16 + 4 * kr + r) * 4u;
      rho[t][r] = bld(rRho, off);
      lam[t][r] = bld(rLam, off);
      mu[t][r] = bld(rMu, off);
      g0[t][r] = bld(rG0, off);
      g1[t][r] = bld(rG1, off);
      g2[t][r] = bld(rG2, off);
    }
  }
  __syncthreads();

#pragma unroll
  for (int gg = 0; gg < 3; ++gg) {
    // ---- 2: t_b = Pq^T g_{3gg+b}, y_b = Vq (A Q)_{3gg+b} (M = cubature points, N = elements,
    // K = nodes); gradient terms; S_b = M(m) t
    {
      const float *Bg = sg + col * RS + 3 * gg * Np + kr;
      const float *Bu = sU + col * RS + 3 * gg * Np + kr;
#pragma unroll
      for (int t = 0; t < L::QTW; ++t) {
        const int qt = wave + 4 * t;
        if (qt >= L::NQT) break;
        f4 aT[KBV], aY[KBV];
#pragma unroll
        for (int kb = 0; kb < KBV; ++kb) {
          aT[kb] = bld_blk<KTV>(rPT, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
          aY[kb] = bld_blk<KTV>(rVq, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
        }
        f4 at[3], ay[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          at[b] = f4{0.f, 0.f, 0.f, 0.f};
          ay[b] = f4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int kb = 0; kb < KBV; ++kb)
#pragma unroll
          for (int sb = 0; sb < 4; ++sb) {
            const int kt = 4 * kb + sb;
            if (kt < KTV) {
#pragma unroll
              for (int b = 0; b < 3; ++b) {
                at[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aT[kb][sb], Bg[b * Np + kt * 4], at[b], 0, 0, 0);
                ay[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aY[kb][sb], Bu[b * Np + kt * 4], ay[b], 0, 0, 0);
              }
            }
          }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = qt * 16 + 4 * kr + r;
          const float ty = at[0][r] * ay[0][r] + at[1][r] * ay[1][r] + at[2][r] * ay[2][r];
          float s0, s1, s2;
          if (gg == 0) {
            g0[t][r] += fdt * ty;
            s0 = rho[t][r] * at[0][r];
            s1 = rho[t][r] * at[1][r];
            s2 = rho[t][r] * at[2][r];
          } else if (gg == 1) {
            const float tt = at[0][r] + at[1][r] + at[2][r];
            g1[t][r] += fdt * tt * (ay[0][r] + ay[1][r] + ay[2][r]);
            g2[t][r] += 2.f * fdt * ty;
            const float m2 = 2.f * mu[t][r], lt = lam[t][r] * tt;
            s0 = m2 * at[0][r] + lt;
            s1 = m2 * at[1][r] + lt;
            s2 = m2 * at[2][r] + lt;
          } else {
            g2[t][r] += fdt * ty;
            s0 = mu[t][r] * at[0][r];
            s1 = mu[t][r] * at[1][r];
            s2 = mu[t][r] * at[2][r];
          }
          const bool in = q < Nq;
          S[q * 16 + col] = in ? s0 : 0.f;
          S[(NQP + q) * 16 + col] = in ? s1 : 0.f;
          S[(2 * NQP + q) * 16 + col] = in ? s2 : 0.f;
        }
      }
    }
    __syncthreads();
    // ---- 3: z_{3gg+b} = Vq^T S_b (M = nodes, K = cubature points) over the group's g rows
    for (int mt = wave; mt < L::NPT; mt += 4) {
      f4 aV[KBP];
#pragma unroll
      for (int kb = 0; kb < KBP; ++kb) aV[kb] = bld_blk<KTP>(rVT, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
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
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aV[kb][sb], Bp[(b * NQP + kt * 4) * 16], acc[b], 0, 0, 0);
          }
        }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int b = 0; b < 3; ++b) sg[col * RS + (3 * gg + b) * Np + n] = acc[b][r];
        }
      }
    }
    __syncthreads();
  }

  // ---- 4: gradient rows (the wave's own cubature tiles, points < Nq), Z rows out; per face
  // node w = 1/2 Fscale LIFT_f^T z (M = face nodes, N = elements, K = nodes) -> (h_v, c) -> H
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t;
    if (qt >= L::NQT) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = qt * 16 + 4 * kr + r;
      if (q < Nq) {
        const unsigned off = (unsigned)(col * Nq + q) * 4u;
        bst(rG0, g0[t][r], off);
        bst(rG1, g1[t][r], off);
        bst(rG2, g2[t][r], off);
      }
    }
  }
  for (int i = tid; i < TB * RowF; i += 256) {
    const int e = i / RowF, rr = i - e * RowF;
    bst(rZ, sg[e * RS + rr], (unsigned)i * 4u);
  }
  {
    const float *zc = sg + col * RS + kr;
    const float tv = md.tau_v, ts = md.tau_s;
    for (int pr = wave; pr < L::NPR; pr += 4) {
      const int f = pr / L::FMT, mt = pr - f * L::FMT;
      f4 aL[KBV];
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb) aL[kb] = bld_blk<KTV>(rLT, lane * 16u, (unsigned)(pr * KBV + kb) * 1024u, kb);
      f4 w[9];
#pragma unroll
      for (int j = 0; j < 9; ++j) w[j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < KTV) {
#pragma unroll
            for (int j = 0; j < 9; ++j)
              w[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(aL[kb][sb], zc[j * Np + kt * 4], w[j], 0, 0, 0);
          }
        }
      const float *Fe = sF + 16 * col + 4 * f;
      const float hs = .5f * Fe[0], nx = Fe[1], ny = Fe[2], nz = Fe[3];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = mt * 16 + 4 * kr + r;
        if (m < Nfp) {
          const float v0 = hs * w[0][r], v1 = hs * w[1][r], v2 = hs * w[2][r];
          const float xx = hs * w[3][r], yy = hs * w[4][r], zz = hs * w[5][r];
          const float yz = hs * w[6][r], xz = hs * w[7][r], xy = hs * w[8][r];
          // b = S^T w_sigma
          const float b0 = xx * nx + xy * ny + xz * nz;
          const float b1 = xy * nx + yy * ny + yz * nz;
          const float b2 = xz * nx + yz * ny + zz * nz;
          // S^T S w_v
          const float s3 = v0 * nx, s4 = v1 * ny, s5 = v2 * nz;
          const float s6 = v2 * ny + v1 * nz, s7 = v2 * nx + v0 * nz, s8 = v1 * nx + v0 * ny;
          const float q0 = s3 * nx + s8 * ny + s7 * nz;
          const float q1 = s8 * nx + s4 * ny + s6 * nz;
          const float q2 = s7 * nx + s6 * ny + s5 * nz;
          float *h = sHt + (col * NfpNf + f * Nfp + m) * 6;
          h[0] = b0 + tv * q0;
          h[1] = b1 + tv * q1;
          h[2] = b2 + tv * q2;
          h[3] = v0 + ts * b0;
          h[4] = v1 + ts * b1;
          h[5] = v2 + ts * b2;
        }
      }
    }
  }
  __syncthreads();
  // ---- 5: H rows out (contiguous)
  for (int i = tid; i < TB * NfpNf * 6; i += 256) bst(rH, sHt[i], (unsigned)i * 4u);
}

// lQ += dt A^T z: flux transpose from own H and the neighbour's H through mapP, volume transpose
// in factored form.
template <int N>
__global__ __launch_bounds__(256) void k_adje2(int K, const float *__restrict__ vgeo,
                                               const float *__restrict__ fgeo,
                                               const int4 *__restrict__ Sids,
                                               const float4 *__restrict__ Swts,
                                               const int4 *__restrict__ Eids,
                                               const float4 *__restrict__ Evals,
                                               const int *__restrict__ FmaskInv,
                                               const int *__restrict__ mapP,
                                               const float *__restrict__ Z,
                                               const float *__restrict__ H, float fdt,
                                               float *__restrict__ lQ) {
  using L = AdjElasLayout<N>;
  constexpr int Np = L::Np, NfpNf = L::NfpNf, Nfp = L::Nfp, TB = L::TB, RowF = L::RowF, Np1 = L::Np1;
  constexpr unsigned ROW = RowF * 4u;
  extern __shared__ float lds[];
  float *sZ = lds;                            // [TB][RowF]
  float *sT = sZ + TB * RowF;                 // [TB][9][Np1]
  float *sSf = sT + TB * 9 * Np1;             // [TB][9][NfpNf]
  float *sG = sSf + TB * 9 * NfpNf;           // [TB][9] in a 256-word slot
  float *sFg = sG + 256;                      // [TB][16]
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * RowF;
  const unsigned tbytes = (unsigned)nk * ROW;
  const brsrc rZ = mkrsrc(Z + base, tbytes), rL = mkrsrc(lQ + base, tbytes);
  const brsrc rHt = mkrsrc(H + (size_t)k0 * NfpNf * 6, (unsigned)nk * NfpNf * 24u);
  const brsrc rHa = mkrsrc(H, (unsigned)K * NfpNf * 24u);
  const brsrc rM = mkrsrc(mapP + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);

  // ---- 1: Z tile and geometry -> LDS
  for (int i = tid; i < TB * RowF; i += 256) sZ[i] = bld(rZ, (unsigned)i * 4u);
  sG[tid] = bld(mkrsrc(vgeo + (size_t)k0 * 9, (unsigned)nk * 36u), tid * 4u);
  sFg[tid] = bld(mkrsrc(fgeo + (size_t)k0 * 16, (unsigned)nk * 64u), tid * 4u);
  __syncthreads();

  // ---- 2: flux transpose per face node: v gets -h_v(own) + h_v(neighbour), sigma gets
  // -S(n)(c(own) + c(neighbour)) (the neighbour's normal is -n); boundary: own terms only
  for (int i = tid; i < TB * NfpNf; i += 256) {
    const int e = i / NfpNf, fn = i - e * NfpNf, f = fn / Nfp;
    const int own = (k0 + e) * NfpNf + fn;
    const int gP = bldi(rM, (unsigned)i * 4u);
    const bool bnd = gP == own || e >= nk;
    float ho[6], hn[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      ho[j] = bld(rHt, (unsigned)(i * 6 + j) * 4u);
      hn[j] = bld(rHa, (unsigned)(gP * 6 + j) * 4u);
    }
    const float m = bnd ? 0.f : 1.f;
    const float *F = sFg + 16 * e + 4 * f;
    const float nx = F[1], ny = F[2], nz = F[3];
    const float a0 = ho[3] + m * hn[3], a1 = ho[4] + m * hn[4], a2 = ho[5] + m * hn[5];
    float *o = sSf + e * 9 * NfpNf + fn;
    o[0] = m * hn[0] - ho[0];
    o[NfpNf] = m * hn[1] - ho[1];
    o[2 * NfpNf] = m * hn[2] - ho[2];
    o[3 * NfpNf] = -a0 * nx;
    o[4 * NfpNf] = -a1 * ny;
    o[5 * NfpNf] = -a2 * nz;
    o[6 * NfpNf] = -(a2 * ny + a1 * nz);
    o[7 * NfpNf] = -(a2 * nx + a0 * nz);
    o[8 * NfpNf] = -(a1 * nx + a0 * ny);
  }
  // ---- 3: T_j = N E^T z_j at the degree N-1 nodes, 9 fields
  for (int i = tid; i < TB * Np1; i += 256) {
    const int e = i / Np1, b = i - e * Np1;
    const int4 s = Sids[b];
    const float4 w = Swts[b];
    const float *z = sZ + e * RowF;
    float *t = sT + e * 9 * Np1 + b;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const float *zj = z + j * Np;
      t[j * Np1] = w.x * zj[s.x] + w.y * zj[s.y] + w.z * zj[s.z] + w.w * zj[s.w];
    }
  }
  __syncthreads();

  // ---- 4: per node: D_m^T z_j = [c_m >= 1] T_j[c - e_m], the transposed strain/divergence
  // couplings of rk_volume_elas, the face terms of the node's faces; lQ += dt A^T z
  for (int i = tid; i < TB * Np; i += 256) {
    const int e = i / Np, c = i - e * Np;
    const int4 E = Eids[c];
    const float4 ev = Evals[c];
    const float mx = ev.x > 0.f ? 1.f : 0.f, my = ev.y > 0.f ? 1.f : 0.f;
    const float mz = ev.z > 0.f ? 1.f : 0.f, mw = ev.w > 0.f ? 1.f : 0.f;
    const float *G = sG + 9 * e;
    float X[9], Yd[9], Zd[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const float *T = sT + (e * 9 + j) * Np1;
      const float u0 = mx * T[E.x], u1 = my * T[E.y], u2 = mz * T[E.z], u3 = mw * T[E.w];
      const float dr = .5f * (u1 - u0), ds = .5f * (u2 - u0), dt = .5f * (u3 - u0);
      X[j] = G[0] * dr + G[3] * ds + G[6] * dt;
      Yd[j] = G[1] * dr + G[4] * ds + G[7] * dt;
      Zd[j] = G[2] * dr + G[5] * ds + G[8] * dt;
    }
    float out[9];
    out[0] = X[3] + Zd[7] + Yd[8];
    out[1] = Yd[4] + Zd[6] + X[8];
    out[2] = Zd[5] + Yd[6] + X[7];
    out[3] = X[0];
    out[4] = Yd[1];
    out[5] = Zd[2];
    out[6] = Zd[1] + Yd[2];
    out[7] = Zd[0] + X[2];
    out[8] = Yd[0] + X[1];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int r = FmaskInv[3 * c + a];
      if (r < 0) continue;
      const float *sf = sSf + e * 9 * NfpNf + r;
#pragma unroll
      for (int j = 0; j < 9; ++j) out[j] += sf[j * NfpNf];
    }
    const unsigned off = (unsigned)(e * RowF + c) * 4u;
    float lq[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) lq[j] = bld(rL, off + j * Np * 4u);
#pragma unroll
    for (int j = 0; j < 9; ++j) bst(rL, lq[j] + fdt * out[j], off + j * Np * 4u);
  }
}

template <int N>
static hipError_t adj_elas_n(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io) {
  using L = AdjElasLayout<N>;
  const int grid = (c->K + L::TB - 1) / L::TB;
  const size_t s1 = sizeof(float) * L::floats1, s2 = sizeof(float) * L::floats2;
  if (s1 > 160 * 1024 || s2 > 160 * 1024) return hipErrorInvalidValue;
  if ((unsigned long long)c->K * 36 * L::Np >= (1ull << 32) ||
      (unsigned long long)c->K * L::NfpNf * 24 >= (1ull << 32))
    return hipErrorInvalidValue;
  hipError_t e;
  if (s1 > 65536 &&
      (e = hipFuncSetAttribute((const void *)k_adje1<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s1)) != hipSuccess)
    return e;
  if (s2 > 65536 &&
      (e = hipFuncSetAttribute((const void *)k_adje2<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s2)) != hipSuccess)
    return e;
  AdjElasMedia md{c->rhoinv, c->lambda, c->mu, c->tau_v, c->tau_s};
  hipLaunchKernelGGL(k_adje1<N>, dim3(grid), dim3(256), s1, c->stream, c->K, fused_ops(c), md, Y, c->fgeo,
                     fa, fb, fdt, lres_io, c->lQ, c->lres, c->Zadj, c->Hadj, c->grad, c->PqTBlk, c->VqTBlk,
                     c->LiftTBlk);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_adje2<N>, dim3(grid), dim3(256), s2, c->stream, c->K, c->vgeo, c->fgeo, c->Sids,
                     c->Swts, c->Eids, c->Evals, c->FmaskInv, c->mapP, c->Zadj, c->Hadj, fdt, c->lQ);
  return hipGetLastError();
}

hipError_t launch_adj_elas_stage(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io) {
  switch (c->N) {
    case 1: return adj_elas_n<1>(c, Y, fa, fb, fdt, lres_io);
    case 2: return adj_elas_n<2>(c, Y, fa, fb, fdt, lres_io);
    case 3: return adj_elas_n<3>(c, Y, fa, fb, fdt, lres_io);
    case 4: return adj_elas_n<4>(c, Y, fa, fb, fdt, lres_io);
    case 5: return adj_elas_n<5>(c, Y, fa, fb, fdt, lres_io);
    case 6: return adj_elas_n<6>(c, Y, fa, fb, fdt, lres_io);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
