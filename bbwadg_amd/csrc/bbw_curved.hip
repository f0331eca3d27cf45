// Curvilinear WADG right-hand side for gfx950: the nodal split-form volume and surface terms
// (rk_volume_WADG_skew_combine, rk_surface_WADG_skew_combine, cpp/okl/WaveKernelsWADG.okl:514-682,
// :3024-3195) for a list of elements in one launch, written to rhs once.
//
// CURVED = true takes the geometric factors at every cubature point (vgeoq, Jq, fgeoq); CURVED =
// false is the same form with the element's constant affine factors (vgeo, J, fgeo), an option for
// the planar elements (BBW_OPT_PLANAR_SPLIT); by default they take k_planar_rhs below, the
// reference's strong form (rk_volume_planar / rk_surface_planar, :2203-2441).  With the degree
// 2N+1 rules both forms are exact integrals of the same terms, so they agree to rounding.
//
// One workgroup (4 waves) owns a tile of TB = 16 listed elements (the N width of the FP32
// 16x16x4 MFMA); A operands are lane-ordered blocks (as bbw_ctx::VqBlk), B operands LDS tiles:
//   1  the tile's nodal state rows -> LDS
//   2  volume: at the cubature points (M = points, K = nodes) u, v, w = Vq (u, v, w) and
//      p_r, p_s, p_t = (Vrq, Vsq, Vtq) p; then per point S_x,y,z = -J grad p and
//      S_r,s,t = J (u . grad r, u . grad s, u . grad t) -> LDS
//   3  rhs_p = Prq S_r + Psq S_s + Ptq S_t, rhs_u,v,w = Pq S_x,y,z into accumulators that stay in
//      registers to the end (M = nodes, K = points)
//   4  surface, FPR faces per round: jumps and averages of the traces at the face nodes -> LDS;
//      interpolated to the face cubature (M = face points, K = face nodes); the fluxes
//      p* = -(avg(u).n - [p]/2) sJ and u* = -([p] - [u].n)/2 sJ n -> LDS; rhs += Pfq flux
//      (M = nodes, K = the round's face points)
//   5  accumulators -> LDS -> coalesced rhs rows
// v_mfma_f32_16x16x4f32 lane layout: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k,
// D[4*(l/16) + r][l%16] in register r of lane l.
#include "bbw_fused_tile.h"

namespace bbw {

template <int N>
struct CurvedLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, Nq = Dims<N>::Nq;
  // points of the degree 2N+1 triangle rule on each face
  static constexpr int Nfq = N == 1 ? 6 : N == 2 ? 7 : N == 3 ? 15 : N == 4 ? 19 : N == 5 ? 28 : 40;
  static constexpr int TB = 16, NW = 4 * Np;
  static constexpr int NQT = (Nq + 15) / 16, NQP = 16 * NQT, NPT = (Np + 15) / 16;
  static constexpr int KTV = (Np + 3) / 4, KBV = (KTV + 3) / 4;      // volume products: K = nodes
  static constexpr int KBP = NQT;                                     // projections: K = NQP points
  static constexpr int NFQT = (Nfq + 15) / 16, NFQP = 16 * NFQT;
  static constexpr int KF = 4 * ((Nfp + 3) / 4), KTF = KF / 4, KBF = (KTF + 3) / 4;  // K = face nodes
  static constexpr int FPR = NFQT >= 4 ? 1 : 4 / NFQT;               // faces per round
  static constexpr int NR = 4 / FPR;
  static_assert(NR * FPR == 4, "rounds cover the four faces");
  static constexpr int NIT = FPR * NFQT;                              // interpolation items per round
  static_assert(NIT <= 4, "one interpolation item per wave");
  static constexpr int RS0 = NW > 3 * Np + 4 * KTV ? NW : 3 * Np + 4 * KTV;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  static constexpr int MTW = (NPT + 3) / 4, QTW = (NQT + 3) / 4;
  static constexpr int SZ = 6 * NQP * 16;              // S [6][NQP][16]
  static constexpr int TZ = 7 * FPR * KF * 16;         // T [7][FPR][KF][16]
  static constexpr int FZ = 4 * FPR * NFQP * 16;       // F [4][FPR * NFQP][16]
  static constexpr int UZ = SZ > TZ + FZ ? SZ : TZ + FZ;
  static constexpr int OU = TB * RS, OEL = OU + UZ;
  static constexpr size_t floats = (size_t)OEL + TB;
};

int curved_nfq(int N) {
  static const int t[7] = {0, 6, 7, 15, 19, 28, 40};
  return N >= 1 && N <= 6 ? t[N] : -1;
}

struct CurvedOps {
  const float *Vq, *Vr, *Vs, *Vt;   // lane-ordered A blocks, M = Nq, K = Np
  const float *Pq, *Pr, *Ps, *Pt;   // M = Np, K = NQP
  const float *Vf;                  // M = Nfq, K = Nfp
  const float *Pf;                  // M = Np, K = 4 NFQP (face f at f NFQP)
  const int *Fmask;
};

template <int N, bool CURVED>
__global__ __launch_bounds__(256) void k_curved_rhs(int Kl, const int *__restrict__ Klist, int K,
                                                    CurvedOps o, const float *__restrict__ vgeoq,
                                                    const float *__restrict__ Jq,
                                                    const float *__restrict__ fgeoq,
                                                    const float *__restrict__ vgeo,
                                                    const float *__restrict__ fgeo,
                                                    const float *__restrict__ Jpl,
                                                    const int *__restrict__ vmapP,
                                                    const float *__restrict__ Q,
                                                    float *__restrict__ rhs) {
  using L = CurvedLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, Nq = L::Nq, Nfq = L::Nfq, TB = L::TB, NW = L::NW;
  constexpr int RS = L::RS, NQP = L::NQP, NFQP = L::NFQP, KF = L::KF, FPR = L::FPR;
  extern __shared__ float lds[];
  float *sQ = lds;                  // [TB][RS]
  float *U = lds + L::OU;           // S, or T | F
  int *sEl = (int *)(lds + L::OEL); // [TB] element ids
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int kk0 = blockIdx.x * TB, nk = min(TB, Kl - kk0);
  // element of each tile column; columns past the list repeat the first (their loads stay in
  // range, their stores are skipped)
  const int kcol = Klist[kk0 + (col < nk ? col : 0)];
  if (tid < TB) sEl[tid] = Klist[kk0 + (tid < nk ? tid : 0)];
  const brsrc rQ = mkrsrc(Q, (unsigned)K * NW * 4u);
  const brsrc rVq = mkrsrc(o.Vq, (unsigned)L::NQT * L::KBV * 1024u);
  const brsrc rVr = mkrsrc(o.Vr, (unsigned)L::NQT * L::KBV * 1024u);
  const brsrc rVs = mkrsrc(o.Vs, (unsigned)L::NQT * L::KBV * 1024u);
  const brsrc rVt = mkrsrc(o.Vt, (unsigned)L::NQT * L::KBV * 1024u);
  const brsrc rPq = mkrsrc(o.Pq, (unsigned)L::NPT * L::KBP * 1024u);
  const brsrc rPr = mkrsrc(o.Pr, (unsigned)L::NPT * L::KBP * 1024u);
  const brsrc rPs = mkrsrc(o.Ps, (unsigned)L::NPT * L::KBP * 1024u);
  const brsrc rPt = mkrsrc(o.Pt, (unsigned)L::NPT * L::KBP * 1024u);
  const brsrc rVf = mkrsrc(o.Vf, (unsigned)L::NFQT * L::KBF * 1024u);
  const brsrc rPf = mkrsrc(o.Pf, (unsigned)L::NPT * 4 * L::NFQT * 1024u);

  // ---- 1: state rows of the tile -> LDS (zero padding past the last field for the K tail)
  {
    constexpr int RR = (NW + 255) / 256;
    float h[RR][TB];
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      const int rr = tid + 256 * k;
#pragma unroll
      for (int e = 0; e < TB; ++e) {
        const int ke = Klist[kk0 + (e < nk ? e : 0)];
        h[k][e] = bld(rQ, rr < NW ? (unsigned)(ke * NW + rr) * 4u : 0x80000000u);
      }
    }
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      const int rr = tid + 256 * k;
      if (rr < NW) {
#pragma unroll
        for (int e = 0; e < TB; ++e) sQ[e * RS + rr] = h[k][e];
      }
    }
    constexpr int KPAD = L::RS0 - NW;
    static_assert(TB * KPAD <= 256, "one padding word per thread");
    if (KPAD > 0 && tid < TB * KPAD) sQ[(tid / KPAD) * RS + NW + tid % KPAD] = 0.f;
  }
  // constant factors of a planar element: rx ry rz sx sy sz tx ty tz, J
  float pg[CURVED ? 1 : 10];
  if constexpr (!CURVED) {
#pragma unroll
    for (int c = 0; c < 9; ++c) pg[c] = vgeo[(size_t)kcol * 9 + c];
    pg[9] = Jpl[kcol];
  }
  __syncthreads();

  // ---- 2: volume values at the cubature points -> S [6][NQP][16]
  float *S = U;
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t;
    if (qt < L::NQT) {
      float gq[CURVED ? 10 : 1][4];
      if constexpr (CURVED) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = min(qt * 16 + 4 * kr + r, Nq - 1);
#pragma unroll
          for (int c = 0; c < 9; ++c) gq[c][r] = vgeoq[((size_t)kcol * 9 + c) * Nq + q];
          gq[9][r] = Jq[(size_t)kcol * Nq + q];
        }
      }
      f4 acc[6];
#pragma unroll
      for (int b = 0; b < 6; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
      const float *Bp = sQ + col * RS + kr;
#pragma unroll
      for (int kb = 0; kb < L::KBV; ++kb) {
        const unsigned ao = (unsigned)(qt * L::KBV + kb) * 1024u;
        const f4 a0 = bld4_s(rVq, lane * 16u, ao), a1 = bld4_s(rVr, lane * 16u, ao);
        const f4 a2 = bld4_s(rVs, lane * 16u, ao), a3 = bld4_s(rVt, lane * 16u, ao);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int kt = 4 * kb + s;
          if (kt < L::KTV) {
            const float bp = Bp[kt * 4], bu = Bp[Np + kt * 4], bv = Bp[2 * Np + kt * 4];
            const float bw = Bp[3 * Np + kt * 4];
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[s], bu, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[s], bv, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[s], bw, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[s], bp, acc[3], 0, 0, 0);
            acc[4] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2[s], bp, acc[4], 0, 0, 0);
            acc[5] = __builtin_amdgcn_mfma_f32_16x16x4f32(a3[s], bp, acc[5], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        float rx, sx, tx, ry, sy, ty, rz, sz, tz, J;
        if constexpr (CURVED) {
          rx = gq[0][r]; sx = gq[1][r]; tx = gq[2][r]; ry = gq[3][r]; sy = gq[4][r];
          ty = gq[5][r]; rz = gq[6][r]; sz = gq[7][r]; tz = gq[8][r]; J = gq[9][r];
        } else {
          rx = pg[0]; ry = pg[1]; rz = pg[2]; sx = pg[3]; sy = pg[4]; sz = pg[5];
          tx = pg[6]; ty = pg[7]; tz = pg[8]; J = pg[9];
        }
        const float u = acc[0][r], v = acc[1][r], w = acc[2][r];
        const float pr = acc[3][r], ps = acc[4][r], pt = acc[5][r];
        const bool in = q < Nq;
        float *Sq = S + q * 16 + col;
        Sq[0] = in ? (u * rx + v * ry + w * rz) * J : 0.f;
        Sq[NQP * 16] = in ? (u * sx + v * sy + w * sz) * J : 0.f;
        Sq[2 * NQP * 16] = in ? (u * tx + v * ty + w * tz) * J : 0.f;
        Sq[3 * NQP * 16] = in ? -(rx * pr + sx * ps + tx * pt) * J : 0.f;
        Sq[4 * NQP * 16] = in ? -(ry * pr + sy * ps + ty * pt) * J : 0.f;
        Sq[5 * NQP * 16] = in ? -(rz * pr + sz * ps + tz * pt) * J : 0.f;
      }
    }
  }
  __syncthreads();

  // ---- 3: rhs_p = Prq S_r + Psq S_s + Ptq S_t; rhs_uvw = Pq S_xyz
  f4 racc[L::MTW][4];
#pragma unroll
  for (int i = 0; i < L::MTW; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) racc[i][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < L::MTW; ++i) {
    const int mt = wave + 4 * i;
    if (mt < L::NPT) {
      const float *Bp = S + kr * 16 + col;
#pragma unroll
      for (int kb = 0; kb < L::KBP; ++kb) {
        const unsigned ao = (unsigned)(mt * L::KBP + kb) * 1024u;
        const f4 aP = bld4_s(rPq, lane * 16u, ao), aR = bld4_s(rPr, lane * 16u, ao);
        const f4 aS = bld4_s(rPs, lane * 16u, ao), aT = bld4_s(rPt, lane * 16u, ao);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int k = (4 * kb + s) * 4;
          racc[i][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(aR[s], Bp[k * 16], racc[i][0], 0, 0, 0);
          racc[i][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(aS[s], Bp[(NQP + k) * 16], racc[i][0], 0, 0, 0);
          racc[i][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(aT[s], Bp[(2 * NQP + k) * 16], racc[i][0], 0, 0, 0);
          racc[i][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[s], Bp[(3 * NQP + k) * 16], racc[i][1], 0, 0, 0);
          racc[i][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[s], Bp[(4 * NQP + k) * 16], racc[i][2], 0, 0, 0);
          racc[i][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[s], Bp[(5 * NQP + k) * 16], racc[i][3], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();

  // ---- 4: surface, FPR faces per round
  float *T = U;                        // [7][FPR][KF][16]
  float *F = U + L::TZ;                // [4][FPR * NFQP][16]
  for (int rd = 0; rd < L::NR; ++rd) {
    // 4a: jumps and averages at the round's face nodes (rows past Nfp are zero)
    for (int p = tid; p < TB * FPR * KF; p += 256) {
      const int e = p / (FPR * KF), rem = p - e * (FPR * KF), fl = rem / KF, n = rem - fl * KF;
      float dp = 0.f, du = 0.f, dv = 0.f, dw = 0.f, au = 0.f, av = 0.f, aw = 0.f;
      if (n < Nfp) {
        const int f = rd * FPR + fl, m = f * Nfp + n, ke = sEl[e];
        const int idM = o.Fmask[m] + ke * NW, idP = vmapP[(size_t)ke * 4 * Nfp + m];
        const float pM = bld(rQ, idM * 4u), uM = bld(rQ, (idM + Np) * 4u);
        const float vM = bld(rQ, (idM + 2 * Np) * 4u), wM = bld(rQ, (idM + 3 * Np) * 4u);
        const float pP = bld(rQ, idP * 4u), uP = bld(rQ, (idP + Np) * 4u);
        const float vP = bld(rQ, (idP + 2 * Np) * 4u), wP = bld(rQ, (idP + 3 * Np) * 4u);
        if (idM == idP) {
          dp = -2.f * pM;
          au = uM; av = vM; aw = wM;
        } else {
          dp = pP - pM; du = uP - uM; dv = vP - vM; dw = wP - wM;
          au = .5f * (uP + uM); av = .5f * (vP + vM); aw = .5f * (wP + wM);
        }
      }
      float *Tp = T + (fl * KF + n) * 16 + e;
      Tp[0] = dp;
      Tp[FPR * KF * 16] = du;
      Tp[2 * FPR * KF * 16] = dv;
      Tp[3 * FPR * KF * 16] = dw;
      Tp[4 * FPR * KF * 16] = au;
      Tp[5 * FPR * KF * 16] = av;
      Tp[6 * FPR * KF * 16] = aw;
    }
    __syncthreads();
    // 4b: to the face cubature points and the fluxes there (one item per wave)
    if (wave < L::NIT) {
      const int fl = wave / L::NFQT, qt = wave - fl * L::NFQT, f = rd * FPR + fl;
      float g[4][4];   // nx, ny, nz, sJ at the lane's four points
      if constexpr (CURVED) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = min(qt * 16 + 4 * kr + r, Nfq - 1);
#pragma unroll
          for (int c = 0; c < 4; ++c) g[c][r] = fgeoq[(((size_t)kcol * 4 + f) * 4 + c) * Nfq + q];
        }
      } else {
        const float *fg = fgeo + (size_t)kcol * 16 + 4 * f;
        const float sJ = fg[0] * pg[9];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          g[0][r] = fg[1]; g[1][r] = fg[2]; g[2][r] = fg[3]; g[3][r] = sJ;
        }
      }
      f4 acc[7];
#pragma unroll
      for (int c = 0; c < 7; ++c) acc[c] = f4{0.f, 0.f, 0.f, 0.f};
      const float *Bp = T + (fl * KF + kr) * 16 + col;
#pragma unroll
      for (int kb = 0; kb < L::KBF; ++kb) {
        const f4 a = bld4_s(rVf, lane * 16u, (unsigned)(qt * L::KBF + kb) * 1024u);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int kt = 4 * kb + s;
          if (kt < L::KTF) {
#pragma unroll
            for (int c = 0; c < 7; ++c)
              acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], Bp[(c * FPR * KF + kt * 4) * 16], acc[c], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        const float nx = g[0][r], ny = g[1][r], nz = g[2][r], sJ = g[3][r];
        const float Unjump = acc[1][r] * nx + acc[2][r] * ny + acc[3][r] * nz;
        const float Unavg = acc[4][r] * nx + acc[5][r] * ny + acc[6][r] * nz;
        const float dp = acc[0][r];
        const float pflux = -(Unavg - .5f * dp) * sJ;
        const float Uflux = -.5f * (dp - Unjump) * sJ;
        const bool in = q < Nfq;
        float *Fp = F + (fl * NFQP + q) * 16 + col;
        Fp[0] = in ? pflux : 0.f;
        Fp[FPR * NFQP * 16] = in ? Uflux * nx : 0.f;
        Fp[2 * FPR * NFQP * 16] = in ? Uflux * ny : 0.f;
        Fp[3 * FPR * NFQP * 16] = in ? Uflux * nz : 0.f;
      }
    }
    __syncthreads();
    // 4c: rhs += Pfq flux over the round's face points
#pragma unroll
    for (int i = 0; i < L::MTW; ++i) {
      const int mt = wave + 4 * i;
      if (mt < L::NPT) {
        const float *Bp = F + kr * 16 + col;
#pragma unroll
        for (int kb = 0; kb < L::NIT; ++kb) {
          const f4 a = bld4_s(rPf, lane * 16u, (unsigned)(mt * 4 * L::NFQT + rd * L::NIT + kb) * 1024u);
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int k = (4 * kb + s) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c)
              racc[i][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], Bp[(c * FPR * NFQP + k) * 16], racc[i][c], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();
  }

  // ---- 5: accumulators -> LDS rows -> rhs
#pragma unroll
  for (int i = 0; i < L::MTW; ++i) {
    const int mt = wave + 4 * i;
    if (mt < L::NPT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int c = 0; c < 4; ++c) sQ[col * RS + c * Np + n] = racc[i][c][r];
        }
      }
    }
  }
  __syncthreads();
  const brsrc rR = mkrsrc(rhs, (unsigned)K * NW * 4u);
  for (int rr = tid; rr < NW; rr += 256) {
#pragma unroll
    for (int e = 0; e < TB; ++e)
      if (e < nk) bst(rR, sQ[e * RS + rr], (unsigned)(sEl[e] * NW + rr) * 4u);
  }
}

// ------------------------------------------------------------------ planar elements
// The reference's strong form for the planar elements of a curved mesh (rk_volume_planar +
// rk_surface_planar, cpp/okl/WaveKernelsWADG.okl:2203-2441), nodal, scaled by J, on MFMA tiles of
// 16 listed elements:
//   1  state rows -> LDS; the upwind fluxes at the face nodes (Fscale J, the face normal) -> LDS
//      F [4][FR][16] (p-flux and the U-flux times nx, ny, nz); the element's constants -> LDS
//   2  u, v, w rows replaced by U.grad r, U.grad s, U.grad t (in place)
//   3  per node tile: div = Dr Ur + Ds Us + Dt Ut, p_r,s,t = (Dr, Ds, Dt) p (M = nodes, K = nodes);
//      rhs = J (-div, -grad p), then rhs += LIFT F (K = face nodes), all in the accumulators
//   4  accumulators -> LDS rows -> rhs
template <int N>
struct PlanarLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf;
  static constexpr int TB = 16, NW = 4 * Np, NPT = (Np + 15) / 16, MTW = (NPT + 3) / 4;
  static constexpr int KTD = (Np + 3) / 4, KBD = (KTD + 3) / 4;       // K = nodes
  static constexpr int KTL = (NfpNf + 3) / 4, KBL = (KTL + 3) / 4;    // K = face nodes
  static constexpr int FR = 4 * KTL;                                  // flux rows (zero tail)
  static constexpr int RS0 = NW > 3 * Np + 4 * KTD ? NW : 3 * Np + 4 * KTD;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  static constexpr int OF = TB * RS, OG = OF + 4 * FR * 16, OEL = OG + TB * 10;
  static constexpr size_t floats = (size_t)OEL + TB;
};

template <int N>
__global__ __launch_bounds__(256) void k_planar_rhs(int Kl, const int *__restrict__ Klist, int K,
                                                    const float *__restrict__ DrB,
                                                    const float *__restrict__ DsB,
                                                    const float *__restrict__ DtB,
                                                    const float *__restrict__ LiftB,
                                                    const int *__restrict__ Fmask,
                                                    const float *__restrict__ vgeo,
                                                    const float *__restrict__ fgeo,
                                                    const float *__restrict__ Jpl,
                                                    const int *__restrict__ vmapP,
                                                    const float *__restrict__ Q,
                                                    float *__restrict__ rhs) {
  using L = PlanarLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, TB = L::TB, NW = L::NW, RS = L::RS;
  constexpr int FR = L::FR;
  extern __shared__ float lds[];
  float *sQ = lds;                    // [TB][RS]
  float *F = lds + L::OF;             // [4][FR][16]
  float *sG = lds + L::OG;            // [TB][10]: rx ry rz sx sy sz tx ty tz J
  int *sEl = (int *)(lds + L::OEL);   // [TB]
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int kk0 = blockIdx.x * TB, nk = min(TB, Kl - kk0);
  const brsrc rQ = mkrsrc(Q, (unsigned)K * NW * 4u);
  const unsigned DB = (unsigned)L::NPT * L::KBD * 1024u;
  const brsrc rDr = mkrsrc(DrB, DB), rDs = mkrsrc(DsB, DB), rDt = mkrsrc(DtB, DB);
  const brsrc rL = mkrsrc(LiftB, (unsigned)L::NPT * L::KBL * 1024u);
  if (tid < TB) sEl[tid] = Klist[kk0 + (tid < nk ? tid : 0)];
  if (tid < TB * 10) {
    const int e = tid / 10, c = tid - e * 10, ke = Klist[kk0 + (e < nk ? e : 0)];
    sG[tid] = c < 9 ? vgeo[(size_t)ke * 9 + c] : Jpl[ke];
  }

  // ---- 1: state rows -> LDS; fluxes at the face nodes -> F
  {
    constexpr int RR = (NW + 255) / 256;
    float h[RR][TB];
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      const int rr = tid + 256 * k;
#pragma unroll
      for (int e = 0; e < TB; ++e) {
        const int ke = Klist[kk0 + (e < nk ? e : 0)];
        h[k][e] = bld(rQ, rr < NW ? (unsigned)(ke * NW + rr) * 4u : 0x80000000u);
      }
    }
    for (int p = tid; p < TB * FR; p += 256) {
      const int e = p / FR, m = p - e * FR;
      float fp = 0.f, fu = 0.f, fv = 0.f, fw = 0.f;
      if (m < NfpNf) {
        const int f = m / Nfp, ke = Klist[kk0 + (e < nk ? e : 0)];
        const int idM = Fmask[m] + ke * NW, idP = vmapP[(size_t)ke * NfpNf + m];
        const float *fg = fgeo + (size_t)ke * 16 + 4 * f;
        const float Fs = fg[0] * Jpl[ke], nx = fg[1], ny = fg[2], nz = fg[3];
        const float pM = bld(rQ, idM * 4u), uM = bld(rQ, (idM + Np) * 4u);
        const float vM = bld(rQ, (idM + 2 * Np) * 4u), wM = bld(rQ, (idM + 3 * Np) * 4u);
        const float pP = bld(rQ, idP * 4u), uP = bld(rQ, (idP + Np) * 4u);
        const float vP = bld(rQ, (idP + 2 * Np) * 4u), wP = bld(rQ, (idP + 3 * Np) * 4u);
        float pj = pP - pM, Uj = (uP - uM) * nx + (vP - vM) * ny + (wP - wM) * nz;
        if (idM == idP) {
          pj = -2.f * pM;
          Uj = 0.f;
        }
        fp = .5f * (pj - Uj) * Fs;
        const float Uf = .5f * (Uj - pj) * Fs;
        fu = Uf * nx; fv = Uf * ny; fw = Uf * nz;
      }
      float *Fp = F + m * 16 + e;
      Fp[0] = fp;
      Fp[FR * 16] = fu;
      Fp[2 * FR * 16] = fv;
      Fp[3 * FR * 16] = fw;
    }
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      const int rr = tid + 256 * k;
      if (rr < NW) {
#pragma unroll
        for (int e = 0; e < TB; ++e) sQ[e * RS + rr] = h[k][e];
      }
    }
    constexpr int KPAD = L::RS0 - NW;
    static_assert(TB * KPAD <= 256, "one padding word per thread");
    if (KPAD > 0 && tid < TB * KPAD) sQ[(tid / KPAD) * RS + NW + tid % KPAD] = 0.f;
  }
  __syncthreads();

  // ---- 2: u, v, w -> U.grad r, U.grad s, U.grad t in place
  for (int p = tid; p < TB * Np; p += 256) {
    const int e = p / Np, n = p - e * Np;
    float *row = sQ + e * RS;
    const float *G = sG + 10 * e;
    const float u = row[Np + n], v = row[2 * Np + n], w = row[3 * Np + n];
    row[Np + n] = u * G[0] + v * G[1] + w * G[2];
    row[2 * Np + n] = u * G[3] + v * G[4] + w * G[5];
    row[3 * Np + n] = u * G[6] + v * G[7] + w * G[8];
  }
  __syncthreads();

  // ---- 3: volume and lift per node tile
  const float *G = sG + 10 * col;
  const float rx = G[0], ry = G[1], rz = G[2], sx = G[3], sy = G[4], sz = G[5];
  const float tx = G[6], ty = G[7], tz = G[8], Jk = G[9];
  f4 o[L::MTW][4];
#pragma unroll
  for (int i = 0; i < L::MTW; ++i) {
    const int mt = wave + 4 * i;
#pragma unroll
    for (int c = 0; c < 4; ++c) o[i][c] = f4{0.f, 0.f, 0.f, 0.f};
    if (mt < L::NPT) {
      f4 ad = f4{0.f, 0.f, 0.f, 0.f}, ar = ad, as = ad, at = ad;
      const float *Bp = sQ + col * RS + kr;
#pragma unroll
      for (int kb = 0; kb < L::KBD; ++kb) {
        const unsigned ao = (unsigned)(mt * L::KBD + kb) * 1024u;
        const f4 aR = bld4_s(rDr, lane * 16u, ao), aS = bld4_s(rDs, lane * 16u, ao);
        const f4 aT = bld4_s(rDt, lane * 16u, ao);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int kt = 4 * kb + s;
          if (kt < L::KTD) {
            const float bp = Bp[kt * 4], b1 = Bp[Np + kt * 4], b2 = Bp[2 * Np + kt * 4];
            const float b3 = Bp[3 * Np + kt * 4];
            ad = __builtin_amdgcn_mfma_f32_16x16x4f32(aR[s], b1, ad, 0, 0, 0);
            ad = __builtin_amdgcn_mfma_f32_16x16x4f32(aS[s], b2, ad, 0, 0, 0);
            ad = __builtin_amdgcn_mfma_f32_16x16x4f32(aT[s], b3, ad, 0, 0, 0);
            ar = __builtin_amdgcn_mfma_f32_16x16x4f32(aR[s], bp, ar, 0, 0, 0);
            as = __builtin_amdgcn_mfma_f32_16x16x4f32(aS[s], bp, as, 0, 0, 0);
            at = __builtin_amdgcn_mfma_f32_16x16x4f32(aT[s], bp, at, 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        o[i][0][r] = -Jk * ad[r];
        o[i][1][r] = -Jk * (rx * ar[r] + sx * as[r] + tx * at[r]);
        o[i][2][r] = -Jk * (ry * ar[r] + sy * as[r] + ty * at[r]);
        o[i][3][r] = -Jk * (rz * ar[r] + sz * as[r] + tz * at[r]);
      }
      const float *Fb = F + kr * 16 + col;
#pragma unroll
      for (int kb = 0; kb < L::KBL; ++kb) {
        const f4 a = bld4_s(rL, lane * 16u, (unsigned)(mt * L::KBL + kb) * 1024u);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int kt = 4 * kb + s;
          if (kt < L::KTL) {
#pragma unroll
            for (int c = 0; c < 4; ++c)
              o[i][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], Fb[(c * FR + kt * 4) * 16], o[i][c], 0, 0, 0);
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- 4: accumulators -> LDS rows -> rhs
#pragma unroll
  for (int i = 0; i < L::MTW; ++i) {
    const int mt = wave + 4 * i;
    if (mt < L::NPT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int c = 0; c < 4; ++c) sQ[col * RS + c * Np + n] = o[i][c][r];
        }
      }
    }
  }
  __syncthreads();
  const brsrc rR = mkrsrc(rhs, (unsigned)K * NW * 4u);
  for (int rr = tid; rr < NW; rr += 256) {
#pragma unroll
    for (int e = 0; e < TB; ++e)
      if (e < nk) bst(rR, sQ[e * RS + rr], (unsigned)(sEl[e] * NW + rr) * 4u);
  }
}

template <int N>
static hipError_t planar_rhs_n(bbw_ctx *c, int Kl, const int *Klist, hipStream_t st) {
  using L = PlanarLayout<N>;
  if (Kl <= 0) return hipSuccess;
  const size_t sh = sizeof(float) * L::floats;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_planar_rhs<N>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_planar_rhs<N>, dim3((Kl + L::TB - 1) / L::TB), dim3(256), sh, st, Kl,
                     Klist, c->K, c->cv_Dr, c->cv_Ds, c->cv_Dt, c->cv_Lift, c->Fmask, c->vgeo,
                     c->fgeo, c->Jplanar, c->vmapP, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

template <int N, bool CURVED>
static hipError_t curved_rhs_n(bbw_ctx *c, int Kl, const int *Klist, hipStream_t st) {
  using L = CurvedLayout<N>;
  if (Kl <= 0) return hipSuccess;
  const size_t sh = sizeof(float) * L::floats;
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_curved_rhs<N, CURVED>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  const CurvedOps o{c->cv_Vq, c->cv_Vr, c->cv_Vs, c->cv_Vt, c->cv_Pq, c->cv_Pr, c->cv_Ps,
                    c->cv_Pt, c->cv_Vf, c->cv_Pf, c->Fmask};
  hipLaunchKernelGGL((k_curved_rhs<N, CURVED>), dim3((Kl + L::TB - 1) / L::TB), dim3(256), sh,
                     st, Kl, Klist, c->K, o, c->vgeoq, c->cv_Jq, c->fgeoq, c->vgeo,
                     c->fgeo, c->Jplanar, c->vmapP, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

// The two element lists write disjoint rhs rows, so the curved list runs on a second stream
// beside the planar one (fork and join through events; inside a graph capture the second stream
// joins the capture): the curved list alone does not fill the GPU.
template <int N>
static hipError_t curved_n(bbw_ctx *c) {
  const bool fork = c->Kcurved > 0 && c->Kplanar > 0 && c->side_stream;
  hipError_t e;
  hipStream_t sc = c->stream;
  if (fork) {
    if ((e = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(c->side_stream, c->ev_fork, 0)) != hipSuccess) return e;
    sc = c->side_stream;
  }
  if ((e = curved_rhs_n<N, true>(c, c->Kcurved, c->Klist_curved, sc)) != hipSuccess) return e;
  // planar elements: the strong form (default) or the split form with constant factors
  e = c->planar_split ? curved_rhs_n<N, false>(c, c->Kplanar, c->Klist_planar, c->stream)
                      : planar_rhs_n<N>(c, c->Kplanar, c->Klist_planar, c->stream);
  if (e != hipSuccess) return e;
  if (fork) {
    if ((e = hipEventRecord(c->ev_join, c->side_stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(c->stream, c->ev_join, 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_curved_rhs(bbw_ctx *c) {
  switch (c->N) {
    case 1: return curved_n<1>(c);
    case 2: return curved_n<2>(c);
    case 3: return curved_n<3>(c);
    case 4: return curved_n<4>(c);
    case 5: return curved_n<5>(c);
    case 6: return curved_n<6>(c);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
