// Discrete adjoint of the fused elastic LSRK45 stage (gfx950): one launch per reverse stage on
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
// A^T z has an element-local part (volume transpose, the own trace's flux transpose) and a
// neighbour part (the flux transpose at the neighbour's trace).  The penalty flux on a face is
// fl_v = S^T dsigma + tau_v S^T S dv, fl_sigma = S (dv + tau_s S^T dsigma), with S = S(n) the
// 3 -> 6 map a -> (a1 nx, a2 ny, a3 nz, a3 ny + a2 nz, a3 nx + a1 nz, a2 nx + a1 ny), so its
// transpose takes w = 1/2 Fscale LIFT_f^T z to h_v = b + tau_v S^T S w_v, h_sigma = S c with
// b = S^T w_sigma, c = w_v + tau_s b: six numbers (h_v, c) per face node.  The own trace gets
// -h, the neighbour's +h (its S(n) is -S(n) of the own face).
// The kernel adds the element-local part to lQ and writes (h_v, c) to Hout; the neighbour part
// is added by the next reverse stage when it loads lQ (from Hin through mapP), so lQ between
// stages lacks the last stage's neighbour terms, which only the next stage reads.  Phases:
//   A  loads: lQ and lres of the tile's (element, node) pairs, the previous stage's (h_v, c) at
//      the neighbour of each face node, media and gradient in the MFMA output layout, geometry
//   B  lQ += dt (pending neighbour terms), g = lres + b lQ -> LDS, lres <- a g; A Q rows -> LDS
//   C  per field group: t = Pq^T g and y = Vq (A Q) on MFMA, gradient terms, S = M t,
//      z = Vq^T S on MFMA
//   D  w = LIFT_f^T z on MFMA -> (h_v, c) -> LDS; T_j = N E^T z_j at the degree N-1 nodes
//   E  per (element, node): D_m^T z_j = [c_m >= 1] T_j[c - e_m], the transposed couplings of
//      rk_volume_elas, the own trace's -h terms; lQ += dt (...); Hout
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
  // KTP: K-steps over the cubature points (S is zero past Nq, so the all-zero steps are skipped)
  static constexpr int KTV = (Np + 3) / 4, KTP = (Nq + 3) / 4, QTW = (NQT + 3) / 4, MTW = (NPT + 3) / 4;
  static constexpr int KBV = (KTV + 3) / 4, KBP = (KTP + 3) / 4;
  static constexpr int FMT = (Nfp + 15) / 16, NPR = 4 * FMT;   // (face, node tile) pairs
  static constexpr int Np1 = N * (N + 1) * (N + 2) / 6;
  // rounds of the thread mappings: (element, node), (element, face node), (element, degree N-1 node)
  static constexpr int RP = (TB * Np + 255) / 256, FR = (TB * NfpNf + 255) / 256;
  static constexpr int BR = (TB * Np1 + 255) / 256;
  // g / z and (A Q) rows [TB][RS]: 9 fields plus the K padding the products read past the
  // ninth, stride 2 mod 32 (MFMA operand reads: 16 rows x 2 consecutive words)
  static constexpr int RS0 = RowF > 8 * Np + 4 * KTV ? RowF : 8 * Np + 4 * KTV;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  // LDS: sg [TB][RS] | U [TB][RS] | S [3][NQP][16] | X | fgeo [TB][16] | vgeo [TB][9] (256) |
  // scratch.  U, S and X hold in turn: the pending neighbour terms [TB][9][NfpNf] (phases A-B);
  // the A Q rows and S (C); H [TB][NfpNf][6] and T [TB][9][Np1] (D-E)
  static constexpr int SFACE = 9 * NfpNf * TB, SHT = 6 * NfpNf * TB, STT = 9 * Np1 * TB;
  static constexpr int RU = TB * RS, RSS = 3 * NQP * 16;
  static constexpr int NEED = SFACE > SHT + STT ? SFACE : SHT + STT;
  static constexpr int RX = NEED > RU + RSS ? NEED - (RU + RSS) : 0;
  static constexpr int OU = TB * RS, OS = OU + RU, OF = OS + RSS + RX, OG = OF + 256, OSC = OG + 256;
  static constexpr size_t floats = (size_t)OSC + 1;
};

struct AdjElasMedia {
  const float *rhoinv, *lambda, *mu;
  float tau_v, tau_s;
};

struct AdjElasMesh {
  const float *vgeo, *fgeo;
  const int *mapP, *FmaskInv;
  const int4 *Sids, *Eids;
  const float4 *Swts, *Evals;
};

template <int N>
__global__ __launch_bounds__(256) void k_adje(int K, FusedOps o, AdjElasMedia md, AdjElasMesh mm,
                                              const float *__restrict__ Y, float fa, float fb, float fdt,
                                              int lres_io, float *__restrict__ lQ, float *__restrict__ lres,
                                              const float *__restrict__ Hin, float *__restrict__ Hout,
                                              float *__restrict__ grad, const float *__restrict__ PqTBlk,
                                              const float *__restrict__ VqTBlk,
                                              const float *__restrict__ LiftTBlk, int xcd_chunk) {
  using L = AdjElasLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, Nq = L::Nq, TB = L::TB, RS = L::RS;
  constexpr int RowF = L::RowF, NQP = L::NQP, KTV = L::KTV, KTP = L::KTP, KBV = L::KBV, KBP = L::KBP;
  constexpr int Np1 = L::Np1, RP = L::RP, FR = L::FR;
  constexpr unsigned ROW = RowF * 4u, NPB = Np * 4u;
  extern __shared__ float lds[];
  float *sg = lds;                       // [TB][RS] g, then z
  float *sU = lds + L::OU;               // [TB][RS] (A Q) of the forward stage
  float *S = lds + L::OS;                // [3][NQP][16]
  float *sFace = lds + L::OU;            // [TB][9][NfpNf] pending neighbour terms (A-B)
  float *sHt = lds + L::OU;              // [TB][NfpNf][6] (h_v, c) (D-E)
  float *sT = lds + L::OU + L::SHT;      // [TB][9][Np1] (D-E)
  float *sF = lds + L::OF;               // [TB][16] fgeo
  float *sG = lds + L::OG;               // [TB][9] vgeo
  float *sScr = lds + L::OSC;            // written by lanes past a row, never read
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = tile_index(xcd_chunk) * TB, nk = min(TB, K - k0);
  const size_t base = (size_t)k0 * RowF, KNq = (size_t)K * Nq;
  const unsigned tbytes = (unsigned)nk * ROW, qbytes = (unsigned)nk * Nq * 4u;
  const brsrc rL = mkrsrc(lQ + base, tbytes), rR = mkrsrc(lres + base, tbytes), rY = mkrsrc(Y + base, tbytes);
  const brsrc rHi = mkrsrc(Hin, (unsigned)K * NfpNf * 24u);
  const brsrc rHo = mkrsrc(Hout + (size_t)k0 * NfpNf * 6, (unsigned)nk * NfpNf * 24u);
  const brsrc rM = mkrsrc(mm.mapP + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);
  const brsrc rFg = mkrsrc(mm.fgeo + (size_t)k0 * 16, (unsigned)nk * 64u);
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
  const bool rd = (lres_io & 1) != 0, st = (lres_io & 2) != 0;

  // ---- A: loads, in the order of first use with the dependent gathers last (the memory counter
  // is in order): the face-node map, then for the (element, node) pairs lQ and lres of the 9
  // fields and the node's face-node slots, normals, geometry, media and gradient, then the
  // gathers through the map.  Lanes past the tile use an out-of-range offset (loads read 0,
  // stores are dropped).
  int gP[FR];
#pragma unroll
  for (int r = 0; r < FR; ++r) {
    const int q = tid + 256 * r, e = q / NfpNf;
    gP[r] = bldi(rM, q < TB * NfpNf && e < nk ? (unsigned)q * 4u : 0x80000000u);
  }
  __builtin_amdgcn_sched_barrier(0);
  float lq[RP][9], lr[RP][9];
  int slot[RP][3];
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int q = tid + 256 * r, e = q / Np, c = q - e * Np;
    const bool in = q < TB * Np;
    const unsigned off = in ? (unsigned)(e * RowF + c) * 4u : 0x80000000u;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      lq[r][j] = bld(rL, off + j * NPB);
      lr[r][j] = bld(rR, rd ? off + j * NPB : 0x80000000u);
    }
    const int cc = in ? c : 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) slot[r][a] = mm.FmaskInv[3 * cc + a];
  }
  // (element, face node) pairs: the own face's normal; below, the previous stage's (h_v, c) at
  // the neighbour's face node (boundary faces and lanes past the tile: no neighbour term)
  float hn[FR][6], nr[FR][3];
#pragma unroll
  for (int r = 0; r < FR; ++r) {
    const int q = tid + 256 * r, e = q / NfpNf, fn = q - e * NfpNf, f = fn / Nfp;
    const bool in = q < TB * NfpNf && e < nk;
#pragma unroll
    for (int d = 0; d < 3; ++d) nr[r][d] = bld(rFg, in ? (unsigned)(e * 16 + 4 * f + 1 + d) * 4u : 0x80000000u);
  }
  const float fgv = bld(rFg, tid * 4u);
  const float vgv = bld(mkrsrc(mm.vgeo + (size_t)k0 * 9, (unsigned)nk * 36u), tid * 4u);
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
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int r = 0; r < FR; ++r)
#pragma unroll
    for (int j = 0; j < 6; ++j) hn[r][j] = bld(rHi, (unsigned)gP[r] * 24u + j * 4u);
  // stores: pending neighbour terms (+h at the neighbour's trace: h_v on v, -S(n) c on sigma
  // with n the own face's normal), geometry
#pragma unroll
  for (int r = 0; r < FR; ++r) {
    const int q = tid + 256 * r, e = q / NfpNf, fn = q - e * NfpNf;
    if (q < TB * NfpNf) {
      const bool inter = e < nk && gP[r] != (k0 + e) * NfpNf + fn;
      const float m = inter ? 1.f : 0.f;
      const float nx = nr[r][0], ny = nr[r][1], nz = nr[r][2];
      const float c0 = m * hn[r][3], c1 = m * hn[r][4], c2 = m * hn[r][5];
      float *o = sFace + e * 9 * NfpNf + fn;
      o[0] = m * hn[r][0];
      o[NfpNf] = m * hn[r][1];
      o[2 * NfpNf] = m * hn[r][2];
      o[3 * NfpNf] = -c0 * nx;
      o[4 * NfpNf] = -c1 * ny;
      o[5 * NfpNf] = -c2 * nz;
      o[6 * NfpNf] = -(c2 * ny + c1 * nz);
      o[7 * NfpNf] = -(c2 * nx + c0 * nz);
      o[8 * NfpNf] = -(c1 * nx + c0 * ny);
    }
  }
  sF[tid] = fgv;
  sG[tid] = vgv;
  __syncthreads();

  // ---- B: lQ += dt (pending neighbour terms at the node's face slots); g = lres + b lQ -> sg;
  // lres <- a g; then the A Q rows and the A operands
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int q = tid + 256 * r, e = q / Np, c = q - e * Np;
    const bool in = q < TB * Np;
    const unsigned off = in ? (unsigned)(e * RowF + c) * 4u : 0x80000000u;
    const float *sf = sFace + (in ? e : 0) * 9 * NfpNf;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      float acc = 0.f;
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (slot[r][a] >= 0) acc += sf[j * NfpNf + slot[r][a]];
      lq[r][j] += fdt * acc;
      const float g = lr[r][j] + fb * lq[r][j];
      if (in) sg[e * RS + j * Np + c] = g;
      if (st) bst(rR, fa * g, off + j * NPB);
    }
  }
  float yv[RP][9];
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int q = tid + 256 * r, e = q / Np, c = q - e * Np;
    const unsigned off = q < TB * Np ? (unsigned)(e * RowF + c) * 4u : 0x80000000u;
#pragma unroll
    for (int j = 0; j < 9; ++j) yv[r][j] = bld(rY, off + j * NPB);
  }
  // A operands used by every group: Pq^T and Vq blocks of the wave's cubature tile, Vq^T blocks
  // of its node tile (when one tile per wave covers them; otherwise loaded per tile below)
  f4 aT[L::QTW == 1 ? KBV : 1], aY[L::QTW == 1 ? KBV : 1], aV[L::MTW == 1 ? KBP : 1];
  if constexpr (L::QTW == 1) {
    const int qt = wave < L::NQT ? wave : L::NQT - 1;
#pragma unroll
    for (int kb = 0; kb < KBV; ++kb) {
      aT[kb] = bld_blk<KTV>(rPT, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
      aY[kb] = bld_blk<KTV>(rVq, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
    }
  }
  if constexpr (L::MTW == 1) {
    const int mt = wave < L::NPT ? wave : L::NPT - 1;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aV[kb] = bld_blk<KTP>(rVT, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int q = tid + 256 * r, e = q / Np, c = q - e * Np;
    float *du = q < TB * Np ? sU + e * RS + c : sScr;
    const int step = q < TB * Np ? Np : 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) du[j * step] = yv[r][j];
  }
  constexpr int KPAD = L::RS0 - RowF;
  static_assert(TB * KPAD <= 256, "one padding word per thread");
  if (KPAD > 0 && tid < TB * KPAD) {
    const int i = (tid / (KPAD > 0 ? KPAD : 1)) * RS + RowF + tid % (KPAD > 0 ? KPAD : 1);
    sg[i] = 0.f;
    sU[i] = 0.f;
  }
  __syncthreads();

#pragma unroll
  for (int gg = 0; gg < 3; ++gg) {
    // ---- C1: t_b = Pq^T g_{3gg+b}, y_b = Vq (A Q)_{3gg+b} (M = cubature points, N = elements,
    // K = nodes); gradient terms; S_b = M(m) t
    {
      const float *Bg = sg + col * RS + 3 * gg * Np + kr;
      const float *Bu = sU + col * RS + 3 * gg * Np + kr;
#pragma unroll
      for (int t = 0; t < L::QTW; ++t) {
        const int qt = wave + 4 * t;
        if (qt >= L::NQT) break;
        f4 bT[KBV], bY[KBV];
#pragma unroll
        for (int kb = 0; kb < KBV; ++kb) {
          if constexpr (L::QTW == 1) {
            bT[kb] = aT[kb];
            bY[kb] = aY[kb];
          } else {
            bT[kb] = bld_blk<KTV>(rPT, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
            bY[kb] = bld_blk<KTV>(rVq, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
          }
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
                at[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bT[kb][sb], Bg[b * Np + kt * 4], at[b], 0, 0, 0);
                ay[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bY[kb][sb], Bu[b * Np + kt * 4], ay[b], 0, 0, 0);
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
    // ---- C2: z_{3gg+b} = Vq^T S_b (M = nodes, K = cubature points) over the group's g rows
    for (int mt = wave; mt < L::NPT; mt += 4) {
      f4 bV[KBP];
#pragma unroll
      for (int kb = 0; kb < KBP; ++kb) {
        if constexpr (L::MTW == 1) bV[kb] = aV[kb];
        else bV[kb] = bld_blk<KTP>(rVT, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
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
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bV[kb][sb], Bp[(b * NQP + kt * 4) * 16], acc[b], 0, 0, 0);
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

  // ---- D: gradient rows (the wave's own cubature tiles, points < Nq); w = 1/2 Fscale
  // LIFT_f^T z (M = face nodes, N = elements, K = nodes) -> (h_v, c) -> sHt; T_j -> sT
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
          float *h = sHt + (col * NfpNf + f * Nfp + m) * 6;
          h[0] = b0 + tv * (s3 * nx + s8 * ny + s7 * nz);
          h[1] = b1 + tv * (s8 * nx + s4 * ny + s6 * nz);
          h[2] = b2 + tv * (s7 * nx + s6 * ny + s5 * nz);
          h[3] = v0 + ts * b0;
          h[4] = v1 + ts * b1;
          h[5] = v2 + ts * b2;
        }
      }
    }
  }
  // the elevation rows of the thread's nodes for phase E
  int4 Er[RP];
  float4 ev[RP];
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int q = tid + 256 * r, c = q - (q / Np) * Np;
    Er[r] = mm.Eids[q < TB * Np ? c : 0];
    ev[r] = mm.Evals[q < TB * Np ? c : 0];
  }
#pragma unroll
  for (int r = 0; r < L::BR; ++r) {
    const int q = tid + 256 * r, e = q / Np1, b = q - e * Np1;
    if (q < TB * Np1) {
      const int4 s = mm.Sids[b];
      const float4 w = mm.Swts[b];
      const float *z = sg + e * RS;
      float *t = sT + e * 9 * Np1 + b;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const float *zj = z + j * Np;
        t[j * Np1] = w.x * zj[s.x] + w.y * zj[s.y] + w.z * zj[s.z] + w.w * zj[s.w];
      }
    }
  }
  __syncthreads();

  // ---- E: per (element, node): volume transpose and the own trace's -h terms; lQ += dt (...)
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int q = tid + 256 * r, e = q / Np, c = q - e * Np;
    if (q >= TB * Np) continue;
    const float mx = ev[r].x > 0.f ? 1.f : 0.f, my = ev[r].y > 0.f ? 1.f : 0.f;
    const float mz = ev[r].z > 0.f ? 1.f : 0.f, mw = ev[r].w > 0.f ? 1.f : 0.f;
    const float *G = sG + 9 * e;
    float dr[9], ds[9], dt3[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const float *T = sT + (e * 9 + j) * Np1;
      const float u0 = mx * T[Er[r].x], u1 = my * T[Er[r].y], u2 = mz * T[Er[r].z], u3 = mw * T[Er[r].w];
      dr[j] = .5f * (u1 - u0);
      ds[j] = .5f * (u2 - u0);
      dt3[j] = .5f * (u3 - u0);
    }
    auto X = [&](int j) { return G[0] * dr[j] + G[3] * ds[j] + G[6] * dt3[j]; };
    auto Yd = [&](int j) { return G[1] * dr[j] + G[4] * ds[j] + G[7] * dt3[j]; };
    auto Zd = [&](int j) { return G[2] * dr[j] + G[5] * ds[j] + G[8] * dt3[j]; };
    float out[9];
    out[0] = X(3) + Zd(7) + Yd(8);
    out[1] = Yd(4) + Zd(6) + X(8);
    out[2] = Zd(5) + Yd(6) + X(7);
    out[3] = X(0);
    out[4] = Yd(1);
    out[5] = Zd(2);
    out[6] = Zd(1) + Yd(2);
    out[7] = Zd(0) + X(2);
    out[8] = Yd(0) + X(1);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int s = slot[r][a];
      if (s < 0) continue;
      const float *h = sHt + (e * NfpNf + s) * 6;
      const float *Fe = sF + 16 * e + 4 * (s / Nfp);
      const float nx = Fe[1], ny = Fe[2], nz = Fe[3], c0 = h[3], c1 = h[4], c2 = h[5];
      out[0] -= h[0];
      out[1] -= h[1];
      out[2] -= h[2];
      out[3] -= c0 * nx;
      out[4] -= c1 * ny;
      out[5] -= c2 * nz;
      out[6] -= c2 * ny + c1 * nz;
      out[7] -= c2 * nx + c0 * nz;
      out[8] -= c1 * nx + c0 * ny;
    }
    const unsigned off = (unsigned)(e * RowF + c) * 4u;
#pragma unroll
    for (int j = 0; j < 9; ++j) bst(rL, lq[r][j] + fdt * out[j], off + j * NPB);
  }
  for (int i = tid; i < TB * NfpNf * 6; i += 256) bst(rHo, sHt[i], (unsigned)i * 4u);
}

template <int N>
static hipError_t adj_elas_n(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io,
                             const float *Hin, float *Hout) {
  using L = AdjElasLayout<N>;
  const int grid = (c->K + L::TB - 1) / L::TB;
  const size_t sh = sizeof(float) * L::floats;
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  if ((unsigned long long)c->K * 36 * L::Np >= (1ull << 32) ||
      (unsigned long long)c->K * L::NfpNf * 24 >= (1ull << 32))
    return hipErrorInvalidValue;
  hipError_t e;
  if (sh > 65536 &&
      (e = hipFuncSetAttribute((const void *)k_adje<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh)) != hipSuccess)
    return e;
  AdjElasMedia md{c->rhoinv, c->lambda, c->mu, c->tau_v, c->tau_s};
  AdjElasMesh mm{c->vgeo, c->fgeo, c->mapP, c->FmaskInv, c->Sids, c->Eids, c->Swts, c->Evals};
  hipLaunchKernelGGL(k_adje<N>, dim3(grid), dim3(256), sh, c->stream, c->K, fused_ops(c), md, mm, Y, fa, fb,
                     fdt, lres_io, c->lQ, c->lres, Hin, Hout, c->grad, c->PqTBlk, c->VqTBlk, c->LiftTBlk,
                     c->xcd_chunk);
  return hipGetLastError();
}

hipError_t launch_adj_elas_stage(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io,
                                 const float *Hin, float *Hout) {
  switch (c->N) {
    case 1: return adj_elas_n<1>(c, Y, fa, fb, fdt, lres_io, Hin, Hout);
    case 2: return adj_elas_n<2>(c, Y, fa, fb, fdt, lres_io, Hin, Hout);
    case 3: return adj_elas_n<3>(c, Y, fa, fb, fdt, lres_io, Hin, Hout);
    case 4: return adj_elas_n<4>(c, Y, fa, fb, fdt, lres_io, Hin, Hout);
    case 5: return adj_elas_n<5>(c, Y, fa, fb, fdt, lres_io, Hin, Hout);
    case 6: return adj_elas_n<6>(c, Y, fa, fb, fdt, lres_io, Hin, Hout);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
