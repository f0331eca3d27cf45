// Fused elastic stage for gfx950: one launch per LSRK stage of the velocity-stress system
// (rk_volume_elas + rk_surface_elas + rk_update_elas, elas_cpp/okl/ElasKernelsWADG.okl:25-433,
// with the Bernstein-Bezier operators), reading Q_in and writing Q_out, so the right-hand side
// never goes through HBM.
//
// One workgroup (4 waves) owns a tile of TB = 16 consecutive elements (the N width of the FP32
// 16x16x4 MFMA).  Fields: v1 v2 v3 sxx syy szz syz sxz sxy.
//   1  loads: Q_in tile -> LDS (kept to the end), geometry, vmapP, then in registers: the
//      neighbour traces of the thread's face nodes, the media at the cubature points (MFMA
//      output layout), the residual rows of the thread's LSRK rows and the shift/elevation
//      table rows
//   2  penalty flux of the 9 fields at the thread's face nodes (Q+ = 0 on the boundary), in
//      registers: the 3 velocity fluxes and the 3-vector w that determines the 6 stress fluxes
//   then, for the groups of three fields whose cubature-point maps are independent (velocities:
//   1/rho; normal stresses: the isotropic stiffness; shear stresses: mu), one of two forms:
//
//   k_fused_elas_cq (BBW_ELAS_LIFT_CUBATURE, the default for N <= 4): the volume elevation and the
//   lift are folded into the cubature operator, Y = Vq (E F + sum_f LIFT_f flux_f) = [Vq E |
//   Vq LIFT_0 | .. | Vq LIFT_3] [F; flux_0; ..; flux_3], one MFMA product per field (K = the
//   degree-(N-1) nodes and the face nodes), so the right-hand side never exists at the nodes before
//   the projection.  One pass for the velocities, one for the six stresses:
//   A  volume step 1: F at the degree-(N-1) nodes b from the shifted coefficients u[b + e_m] of the
//      fields the outputs differentiate; the face values (velocity fluxes z, or the 3-vector w of
//      the stress fluxes) -> LDS as B operand rows
//   B  Y on MFMA (M = cubature points).  The stresses' flux part is formed per face f and component
//      c as P = [Vq LIFT_f] w_c, which enters three stresses with the face's normal (sxx += nx P_x,
//      sxz += nz P_x, sxy += ny P_x, ...), so w is lifted once for both stress groups
//   then per group: S = map_g(Y) -> LDS; C rhs = Pq S on MFMA (M = nodes); D LSRK for the group's
//   rows: res = fa res + fdt (rhs + source); Q_out = Q_in + fb res
//
//   k_fused_elas (BBW_ELAS_LIFT_NODES, and the adjoint driver's forward passes, which keep A Q at
//   the nodes): per group, the right-hand side is formed at the nodes, then projected:
//   3  volume step 1 (as A); the group's flux -> LDS as MFMA B operands
//   4  volume step 2: rhs_g[a] = sum_j a_j F_g[a - e_j] (elevation) -> LDS
//   5  lift on MFMA: rhs_g += LIFT_f flux_f per face (M = nodes, N = elements, K = face nodes);
//      the six stress fluxes are the symmetric part of n (x) w with one 3-vector w per face node
//      and the normal constant on a face, so the stresses lift w (3 products per face, shared by
//      both stress groups) and apply the normals afterwards
//   6  S = map_g(Vq rhs_g) on MFMA (M = cubature points, K = nodes)
//   7  rhs_g = Pq S on MFMA (M = nodes, K = cubature points)
//   8  LSRK as D
// The derivative in factored form: D_m u[a] = sum_j a_j u[a - e_j + e_m], and every output is
// linear in the D_m u, so the volume terms are an elevation of values formed at the degree-(N-1)
// nodes.  The lift is LIFT = EEL blkdiag(L0) applied densely (as rk_surface_elas lifts densely).
//
// v_mfma_f32_16x16x4f32 lane layout: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k,
// D[4*(l/16) + r][l%16] in register r of lane l.
#include "bbw_fused_tile.h"

namespace bbw {

// Thread mappings and sizes shared by both forms of the stage.
template <int N>
struct ElasTileLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, Nq = Dims<N>::Nq;
  static constexpr int TB = 16, RowF = 9 * Np;
  static constexpr int NQT = (Nq + 15) / 16, NQP = NQT * 16, NPT = (Np + 15) / 16;
  // KTP: K-steps over the cubature points (S is zero past Nq, so the all-zero steps are skipped)
  static constexpr int KTP = (Nq + 3) / 4, QTW = (NQT + 3) / 4, KBP = (KTP + 3) / 4;
  static constexpr int KF = 4 * ((Nfp + 3) / 4), KTL = KF / 4, KBL = (KTL + 3) / 4;
  static constexpr int MTW = (NPT + 3) / 4;   // node tiles per wave in the products with M = nodes
  static constexpr int NG3 = 256 / NfpNf, EPT3 = (TB + NG3 - 1) / NG3;   // face nodes
  static constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS, EPTV = (TB + NG - 1) / NG;  // nodes
  static constexpr int Np1 = N * (N + 1) * (N + 2) / 6;
  static constexpr int GS1 = Np1 <= 32 ? 32 : 64, NG1 = 256 / GS1, EPT1 = TB / NG1;   // degree N-1 nodes
  static_assert(Np1 <= 64 && NG1 * EPT1 == TB, "volume step 1 covers the tile");
  // LSRK rows of a group: 3 Np rows x TB elements; HE element halves when 2 * 3 Np <= 256
  static constexpr int LR = 3 * Np, HE = 2 * LR <= 256 ? 2 : 1, EPL = TB / HE;
  static_assert(LR <= 256, "one LSRK row per thread");
  static constexpr bool PFR = HE == 2;   // residual rows prefetched into registers
  // the Q_in tile padded to whole 4 KB rounds of 256 threads x 16 bytes
  static constexpr unsigned QB = TB * RowF * 4u, QIT = (QB + 4095) / 4096;
  static constexpr int QTP = (int)QIT * 1024;
  // fgeo rows in LDS [TB][FGS]: the odd stride spreads the per-element normal reads of a wave
  // (16 elements, one per lane column) over the banks
  static constexpr int FGS = 17, GW = TB * 9 + TB * FGS;
};

template <int N>
struct ElasFusedLayout : ElasTileLayout<N> {
  using B = ElasTileLayout<N>;
  static constexpr int KTV = (B::Np + 3) / 4, KBV = (KTV + 3) / 4;
  // a group's rhs rows: 3 fields of Np plus the K padding the Vq product reads past the third,
  // row stride rounded up to 2 mod 32 (MFMA operand reads: 16 rows x 2 consecutive words)
  static constexpr int RG0 = 3 * B::Np > 2 * B::Np + 4 * KTV ? 3 * B::Np : 2 * B::Np + 4 * KTV;
  static constexpr int RG = RG0 + ((34 - RG0 % 32) % 32);
  static constexpr int KB = B::NfpNf + 4, BS = B::TB + 1;  // flux rows (zero tail) and their stride
  // LDS: Q_in tile | geometry vgeo [TB][9], fgeo [TB][16] | F float4 [TB][Np1] | rhs_g [TB][RG] |
  // W: flux [3][KB][BS] or S [3][NQP][16] | slots [TB]
  static constexpr int OQ = 0, OG = B::QTP, OF = OG + B::GW, OR = OF + 4 * B::TB * B::Np1;
  static constexpr int WSZ = 3 * KB * BS > 3 * B::NQP * 16 ? 3 * KB * BS : 3 * B::NQP * 16;
  static constexpr int OW = OR + B::TB * RG, OSL = OW + WSZ;
  static constexpr size_t floats = (size_t)OSL + B::TB;
  static_assert(3 * (KB - B::NfpNf) * BS <= 256, "one flux tail word per thread");
  static_assert(B::TB * (RG0 - 3 * B::Np) <= 256, "one rhs padding word per thread");
};

template <int N>
struct ElasCqLayout : ElasTileLayout<N> {
  using B = ElasTileLayout<N>;
  static constexpr int KT1 = (B::Np1 + 3) / 4;                         // K-steps over F
  static constexpr int KTC = KT1 + 4 * B::KTL, KBC = (KTC + 3) / 4;    // the folded product
  // B operand rows (stride BS = 17, which spreads column writes over banks).  F of a field:
  // rows [0, Np1), zero up to R1 = 4 KT1.  Face values of a field: face node m (face-major, Nfp
  // per face) at row m, and a zero tail of 4 rows (the last face's last K-step reads past its Nfp
  // rows, against zero columns of the operator).
  //   velocity pass: per field b, [F (R1 rows) | face values (RT rows)] at b RBV
  //   stress pass:   F of the 6 stresses at b R1, then the 3 components of w at 6 R1 + c RT
  static constexpr int R1 = 4 * KT1, RT = B::NfpNf + 4, RBV = R1 + RT, BS = B::TB + 1;
  // S transposed, [3][TB][SQS] (element rows of the cubature values): the product's four
  // consecutive points per lane go out in one 128-bit store, and SQS = 4 mod 32 spreads the Pq
  // operand reads (16 elements x 4 consecutive points) over the banks
  static constexpr int SQS = B::NQP + 4;
  static constexpr int WV = 3 * RBV * BS, WT = (6 * R1 + 3 * RT) * BS, WS = 3 * B::TB * SQS;
  static constexpr int W = WT > WV ? (WT > WS ? WT : WS) : (WV > WS ? WV : WS);
  static constexpr int NZF = R1 - B::Np1;                               // zero F rows per field
  static constexpr int NZV = 3 * (NZF + 4) * BS, NZT = (6 * NZF + 3 * 4) * BS;   // zero words
  // rhs rows [TB][RG] after the Pq product (odd stride: its column stores spread over banks)
  static constexpr int RG = (3 * B::Np) | 1;
  // LDS: Q_in tile | geometry | W: B operands, then S [3][TB][SQS] | rhs | slots
  static constexpr int OQ = 0, OG = B::QTP, OW = OG + B::GW, OR = OW + W, OSL = OR + B::TB * RG;
  static constexpr size_t floats = (size_t)OSL + B::TB;
};

struct ElasMedia {
  const float *rhoinv, *lambda, *mu;
  float tau_v, tau_s;
};

// A tile's register inputs (phase 1) and, after elas_flux, its face fluxes.
template <int N>
struct ElasTileIn {
  using L = ElasTileLayout<N>;
  int fid3;                  // the thread's face node (flux phase): its node id in the element
  int vm[L::EPT3];           // vmapP of the thread's face nodes
  float4 qreg[L::QIT];       // the Q_in tile on its way to LDS
  float gv0, gf0;            // one geometry word each (vgeo, fgeo)
  int4 Sr, Er;               // shift / elevation table rows
  float4 evr;
  float rho[L::QTW][4], lam[L::QTW][4], mu[L::QTW][4];   // media in the MFMA output layout
  float rreg[L::PFR ? 3 : 1][L::PFR ? L::EPL : 1];       // residual rows (with PFR)
  float fl[L::EPT3][9];      // neighbour traces, then the fluxes (elas_flux)
};

// Phase 1 loads: vmapP first (the trace gather waits for it), the Q_in tile and geometry, the
// table rows, media, residual rows, then the traces.  Requires K * 9 * Np * 4 < 2^32.
template <int N>
__device__ __forceinline__ void elas_tile_load(ElasTileIn<N> &in, const FusedMesh &m, const FusedOps &o,
                                               const ElasMedia &md, const float *Qin, const float *res,
                                               float fa, int k0, int nk, const int tid) {
  using L = ElasTileLayout<N>;
  constexpr int Np = L::Np, NfpNf = L::NfpNf, Nq = L::Nq, RowF = L::RowF;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const unsigned tbytes = (unsigned)nk * RowF * 4u;
  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf;
  in.fid3 = o.Fmask[m3 < NfpNf ? m3 : 0];
  {
    const brsrc rv = mkrsrc(m.vmapP + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) in.vm[i] = bldi(rv, (unsigned)((g3 + L::NG3 * i) * NfpNf + m3) * 4u);
  }
  {
    const brsrc rQ = mkrsrc(Qin + (size_t)k0 * RowF, tbytes);
#pragma unroll
    for (unsigned it = 0; it < L::QIT; ++it) in.qreg[it] = bld4(rQ, tid * 16u + it * 4096u);
  }
  in.gv0 = bld(mkrsrc(m.vgeo + (size_t)k0 * 9, (unsigned)nk * 36u), tid * 4u);
  in.gf0 = bld(mkrsrc(m.fgeo + (size_t)k0 * 16, (unsigned)nk * 64u), tid * 4u);
  const int nv = tid % L::GS, n1 = tid % L::GS1;
  in.Sr = o.Sids[n1 < L::Np1 ? n1 : 0];
  in.Er = o.Eids[nv < Np ? nv : 0];
  in.evr = o.Evals[nv < Np ? nv : 0];
  // rows q >= Nq read a neighbouring value; the padded Vq rows make them 0
  {
    const unsigned qbytes = (unsigned)nk * Nq * 4u;
    const brsrc rRho = mkrsrc(md.rhoinv + (size_t)k0 * Nq, qbytes);
    const brsrc rLam = mkrsrc(md.lambda + (size_t)k0 * Nq, qbytes);
    const brsrc rMu = mkrsrc(md.mu + (size_t)k0 * Nq, qbytes);
    const unsigned co = (unsigned)(col * Nq + 4 * kr) * 4u;
#pragma unroll
    for (int t = 0; t < L::QTW; ++t) {
      const int qt = wave + 4 * t < L::NQT ? wave + 4 * t : L::NQT - 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        in.rho[t][r] = bld(rRho, co + (qt * 16 + r) * 4u);
        in.lam[t][r] = bld(rLam, co + (qt * 16 + r) * 4u);
        in.mu[t][r] = bld(rMu, co + (qt * 16 + r) * 4u);
      }
    }
  }
  // LSRK rows: row lr of the group (field 3g + lr / Np, node lr % Np), elements he * EPL ..
  if constexpr (L::PFR) {
    const int lr = tid % L::LR, he = tid / L::LR;
    const bool onl = he < L::HE;
    const brsrc rRes = mkrsrc(res + (size_t)k0 * RowF, tbytes);
#pragma unroll
    for (int g = 0; g < 3; ++g)
#pragma unroll
      for (int i = 0; i < L::EPL; ++i) {
        const unsigned off = onl ? (unsigned)((he * L::EPL + i) * RowF + 3 * g * Np + lr) * 4u : 0x80000000u;
        in.rreg[g][i] = fa != 0.f ? bld(rRes, off) : 0.f;
      }
  }
  // neighbour traces of the 9 fields (then the flux, in the same registers)
  {
    const brsrc rT = mkrsrc(Qin, (unsigned)m.K * RowF * 4u);
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i)
#pragma unroll
      for (int c = 0; c < 9; ++c) in.fl[i][c] = bld(rT, (unsigned)(in.vm[i] + c * Np) * 4u);
  }
}

template <int N>
__device__ __forceinline__ void elas_tile_store(const ElasTileIn<N> &in, float *sQ, float *sG, float *sFg,
                                                const int tid) {
  using L = ElasTileLayout<N>;
  float4 *sQ4 = (float4 *)sQ;
#pragma unroll
  for (unsigned it = 0; it < L::QIT; ++it) sQ4[tid + it * 256u] = in.qreg[it];
  if (tid < L::TB * 9) sG[tid] = in.gv0;
  sFg[(tid >> 4) * L::FGS + (tid & 15)] = in.gf0;
}

// Phase 2: penalty flux at the face nodes (central flux + tau_v / tau_s penalties, Q+ = 0 on
// the boundary), the same formula as rk_surface_elas: fl[i][0..2] the velocity fluxes,
// fl[i][3..5] the 3-vector w of the stress fluxes.
template <int N>
__device__ __forceinline__ void elas_flux(ElasTileIn<N> &in, const float *sQ, const float *sFg,
                                          const ElasMedia &md, int k0, const int tid) {
  using L = ElasTileLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, TB = L::TB, RowF = L::RowF;
  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf, f3 = m3 / Nfp;
  if (g3 < L::NG3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) {
        const float *q = sQ + e * RowF + in.fid3;
        const float *F = sFg + L::FGS * e + 4 * f3;
        const float Fs = .5f * F[0], nx = F[1], ny = F[2], nz = F[3];
        const bool isB = in.vm[i] == in.fid3 + (k0 + e) * RowF;
        float dQ[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) dQ[c] = (isB ? 0.f : in.fl[i][c]) - q[c * Np];
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
        // velocity fluxes z = Fs (sigma jump . n + tau_v A_n^T A_n v jump)
        in.fl[i][0] = Fs * (fc[0] + md.tau_v * (fc[3] * nx + fc[8] * ny + fc[7] * nz));
        in.fl[i][1] = Fs * (fc[1] + md.tau_v * (fc[8] * nx + fc[4] * ny + fc[6] * nz));
        in.fl[i][2] = Fs * (fc[2] + md.tau_v * (fc[7] * nx + fc[6] * ny + fc[5] * nz));
        // stress fluxes: the symmetric part of n (x) w with w = Fs (v jump + tau_s sigma jump . n)
        // (sxx: nx w_x, syz: ny w_z + nz w_y, ...); only w is kept, the normals are applied later
        in.fl[i][3] = Fs * (dQ[0] + md.tau_s * fc[0]);
        in.fl[i][4] = Fs * (dQ[1] + md.tau_s * fc[1]);
        in.fl[i][5] = Fs * (dQ[2] + md.tau_s * fc[2]);
      }
    }
  }
}

// Volume step 1 of group g (a constant after unrolling) at the thread's degree-(N-1) node for
// element e: the values F whose elevation is the group's volume term, from the shifted
// coefficients u[b + e_m] of the Q_in tile and the element's geometric factors.  g = 0: F[0..2]
// = div sigma; g = 1, 2: F[0..2] the normal and F[3..5] the engineering shear strain rates of v
// (both stress groups differentiate the same three fields).
template <int N>
__device__ __forceinline__ void elas_vol1(int g, const float *q, const float *G, const int4 Sr, float (&F)[6]) {
  constexpr int Np = ElasTileLayout<N>::Np;
  // gradients from (d/dr, d/ds, d/dt) = ((u[b+e2] - u[b+e1]) / 2, (u[b+e3] - u[b+e1]) / 2,
  // (u[b+e4] - u[b+e1]) / 2)
  auto grad = [&](int c, float &dx, float &dy, float &dz) {
    const float *qc = q + c * Np;
    const float s0 = qc[Sr.x], s1 = qc[Sr.y], s2 = qc[Sr.z], s3 = qc[Sr.w];
    const float dr = .5f * (s1 - s0), ds = .5f * (s2 - s0), dt = .5f * (s3 - s0);
    dx = G[0] * dr + G[3] * ds + G[6] * dt;
    dy = G[1] * dr + G[4] * ds + G[7] * dt;
    dz = G[2] * dr + G[5] * ds + G[8] * dt;
  };
  if (g == 0) {
    // div sigma: (sxx_x + sxy_y + sxz_z, sxy_x + syy_y + syz_z, sxz_x + syz_y + szz_z)
    float x3, y3, z3, x4, y4, z4, x5, y5, z5, x6, y6, z6, x7, y7, z7, x8, y8, z8;
    grad(3, x3, y3, z3);
    grad(4, x4, y4, z4);
    grad(5, x5, y5, z5);
    grad(6, x6, y6, z6);
    grad(7, x7, y7, z7);
    grad(8, x8, y8, z8);
    F[0] = x3 + y8 + z7;
    F[1] = x8 + y4 + z6;
    F[2] = x7 + y6 + z5;
  } else {
    float x0, y0, z0, x1, y1, z1, x2, y2, z2;
    grad(0, x0, y0, z0);
    grad(1, x1, y1, z1);
    grad(2, x2, y2, z2);
    F[0] = x0;        // normal strain rates
    F[1] = y1;
    F[2] = z2;
    F[3] = y2 + z1;   // engineering shear strain rates (yz, xz, xy)
    F[4] = x2 + z0;
    F[5] = x1 + y0;
  }
}

// Phase D / 8: the LSRK stage for group g's rows from the projected rhs rows sR [TB][RG]:
// res = fa res + fdt (rhs + source); Q_out = Q_in + fb res.
template <int N, int RG>
__device__ __forceinline__ void elas_lsrk(int g, const ElasTileIn<N> &in, const float *sR, const float *sQ,
                                          const int *sSlot, const SrcArgs &src, bool has_src, float ft,
                                          float fa, float fb, float fdt, int store_res, brsrc rRes, brsrc rOut,
                                          const int tid) {
  using L = ElasTileLayout<N>;
  constexpr int Np = L::Np, RowF = L::RowF;
  const int lr = tid % L::LR, he = tid / L::LR;
  if (he < L::HE) {
    const int fr = 3 * g + lr / Np, nr = lr - (lr / Np) * Np;
    float wsel = 0.f;
#pragma unroll
    for (int f = 0; f < 9; ++f) wsel = fr == f ? src.w[f] : wsel;
    float rv[L::EPL];
#pragma unroll
    for (int i = 0; i < L::EPL; ++i) {
      const unsigned off = (unsigned)((he * L::EPL + i) * RowF + 3 * g * Np + lr) * 4u;
      if constexpr (L::PFR) rv[i] = in.rreg[g][i];
      else rv[i] = fa != 0.f ? bld(rRes, off) : 0.f;
    }
#pragma unroll
    for (int i = 0; i < L::EPL; ++i) {
      const int e = he * L::EPL + i;
      float v = sR[e * RG + lr];
      if (has_src) {
        const int slot = sSlot[e];
        if (slot >= 0) v += wsel * ft * src.coef[(size_t)slot * Np + nr];
      }
      const float x = fa * rv[i] + fdt * v;
      const unsigned off = (unsigned)(e * RowF + 3 * g * Np + lr) * 4u;
      bst(rOut, sQ[e * RowF + 3 * g * Np + lr] + fb * x, off);
      if (store_res) bst(rRes, x, off);
    }
  }
}

// CAP: also write the right-hand side A Q_in (before the WADG projection, without the source)
// to yout [K][9][Np], the stage value the discrete adjoint's media gradient needs.
template <int N, bool CAP>
__global__ __launch_bounds__(256, N <= 4 ? 3 : N == 5 ? 2 : 1) void k_fused_elas(FusedMesh m, FusedOps o, ElasMedia md, const float *Qin,
                                                        float *Qout, float *res, SrcArgs src, float fa, float fb,
                                                        float fdt, float ft, const float *ftime_dev, int store_res,
                                                        float *yout) {
  using L = ElasFusedLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, Nq = L::Nq, TB = L::TB, RowF = L::RowF;
  constexpr int RG = L::RG, KB = L::KB, BS = L::BS, NQP = L::NQP, NPT = L::NPT, Np1 = L::Np1;
  constexpr int KTV = L::KTV, KTP = L::KTP, KBV = L::KBV, KBP = L::KBP, KTL = L::KTL, KBL = L::KBL;
  extern __shared__ float lds[];
  float *sQ = lds + L::OQ;                 // [TB][RowF]
  float *sG = lds + L::OG;                 // vgeo [TB][9]
  float *sFg = sG + TB * 9;                // fgeo [TB][FGS]
  float4 *sF = (float4 *)(lds + L::OF);    // [TB][Np1]
  float *sR = lds + L::OR;                 // [TB][RG]
  float *sB = lds + L::OW;                 // [3][KB][BS]  (phases 3-5)
  float *S = lds + L::OW;                  // [3][NQP][16] (phases 6-7)
  int *sSlot = (int *)(lds + L::OSL);
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = tile_index(m.xcd_chunk) * TB, nk = min(TB, m.K - k0);
  const unsigned tbytes = (unsigned)nk * RowF * 4u;
  if (ftime_dev) ft = *ftime_dev;
  static_assert(TB == 16, "the per-tile source table is built for 16-element tiles");
  const int2 stile = src_tile(src, k0, nk);
  const bool has_src = stile.y != 0;

  // ---- 1: loads, then the LDS stores
  ElasTileIn<N> in;
  elas_tile_load<N>(in, m, o, md, Qin, res, fa, k0, nk, tid);
  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf;
  const bool on3 = g3 < L::NG3;
  const int gv = tid / L::GS, nv = tid - gv * L::GS;
  const bool onv = gv < L::NG && nv < Np;
  const int g1 = tid / L::GS1, n1 = tid - g1 * L::GS1;
  const bool on1 = n1 < Np1;
  const int4 Sr = in.Sr, Er = in.Er;
  const float4 evr = in.evr;
  const auto &rho = in.rho, &lam = in.lam, &mu = in.mu;
  const auto &fl = in.fl;
  const brsrc rRes = mkrsrc(res + (size_t)k0 * RowF, tbytes);
  const brsrc rOut = mkrsrc(Qout + (size_t)k0 * RowF, tbytes);
  elas_tile_store<N>(in, sQ, sG, sFg, tid);
  // zero K padding of the rhs rows (read by the Vq product, never written otherwise)
  constexpr int KPAD = L::RG0 - 3 * Np;
  if (KPAD > 0 && tid < TB * KPAD) sR[(tid / KPAD) * RG + 3 * Np + tid % KPAD] = 0.f;
  if (has_src && tid < TB) sSlot[tid] = src_slot_tile(src, stile, k0, tid);
  __syncthreads();

  // ---- 2: penalty flux at the face nodes
  elas_flux<N>(in, sQ, sFg, md, k0, tid);

  // MFMA A operands: lift blocks of the wave's first node tile (every group uses them), the Vq
  // blocks of the wave's cubature tiles and the Pq blocks of its first node tile.  The loads
  // are unconditional (clamped to the last tile): loads under a wave test leave the compiler
  // unsure how many are in flight where the paths join.
  const brsrc rAL = mkrsrc(o.LiftBlk, 4u * NPT * KBL * 1024u);
  const brsrc rAV = mkrsrc(o.VqBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rAP = mkrsrc(o.PqBlk, (unsigned)NPT * KBP * 1024u);
  const int mt0 = wave < NPT ? wave : NPT - 1;
  f4 aL[4][KBL], aV[L::QTW][KBV], aP[KBP];
  auto load_lift = [&](int mt) {
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int b = 0; b < KBL; ++b)
        aL[f][b] = bld_blk<KTL>(rAL, lane * 16u, (unsigned)((f * NPT + mt) * KBL + b) * 1024u, b);
  };
  auto load_pq = [&](int mt) {
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<KTP>(rAP, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
  };
  load_lift(mt0);
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t < L::NQT ? wave + 4 * t : L::NQT - 1;
#pragma unroll
    for (int kb = 0; kb < KBV; ++kb) aV[t][kb] = bld_blk<KTV>(rAV, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
  }
  load_pq(mt0);

  float shr[L::MTW][3][4];   // shear-stress lift sums of the wave's node tiles (lift phase)
  // unrolled: the group selects flux registers and media maps at compile time
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    // ---- 3: volume step 1 (F_g) and the group's flux as MFMA B operands
    if (on1) {
#pragma unroll
      for (int i = 0; i < L::EPT1; ++i) {
        const int e = g1 + L::NG1 * i;
        float Fv[6];
        elas_vol1<N>(g, sQ + e * RowF, sG + 9 * e, Sr, Fv);
        const float F0 = g == 2 ? Fv[3] : Fv[0], F1 = g == 2 ? Fv[4] : Fv[1], F2 = g == 2 ? Fv[5] : Fv[2];
        sF[e * Np1 + n1] = make_float4(F0, F1, F2, 0.f);
      }
    }
    if (g < 2 && on3) {
#pragma unroll
      for (int i = 0; i < L::EPT3; ++i) {
        const int e = g3 + L::NG3 * i;
        if (L::NG3 * L::EPT3 == TB || e < TB) {
#pragma unroll
          for (int b = 0; b < 3; ++b) sB[(b * KB + m3) * BS + e] = fl[i][3 * g + b];
        }
      }
    }
    if (g < 2 && tid < 3 * (KB - NfpNf) * BS) {
      const int b = tid / ((KB - NfpNf) * BS), j = tid - b * ((KB - NfpNf) * BS);
      sB[(b * KB + NfpNf) * BS + j] = 0.f;
    }
    __syncthreads();

    // ---- 4: volume step 2 (elevation) -> rhs_g
    if (onv) {
#pragma unroll
      for (int i = 0; i < L::EPTV; ++i) {
        const int e = gv + L::NG * i;
        if (L::NG * L::EPTV == TB || e < TB) {
          const float4 *Fb = sF + e * Np1;
          const float4 f1 = Fb[Er.x], f2 = Fb[Er.y], f3v = Fb[Er.z], f4v = Fb[Er.w];
          float *r = sR + e * RG + nv;
          r[0] = evr.x * f1.x + evr.y * f2.x + evr.z * f3v.x + evr.w * f4v.x;
          r[Np] = evr.x * f1.y + evr.y * f2.y + evr.z * f3v.y + evr.w * f4v.y;
          r[2 * Np] = evr.x * f1.z + evr.y * f2.z + evr.z * f3v.z + evr.w * f4v.z;
        }
      }
    }
    __syncthreads();

    // ---- 5: lift on MFMA (M = nodes, N = elements, K = face nodes).  Velocities:
    // rhs_c += sum_f LIFT_f z_{c,f}.  Stresses: the per-face products P_{f,c} = LIFT_f w_{c,f} are
    // formed once (in the normal-stress group) and combined with the face normals:
    // sxx += sum_f nx_f P_{f,x}, ..., syz += sum_f (ny_f P_{f,z} + nz_f P_{f,y}), ...; the shear
    // sums wait in registers for the shear group.
    {
      const float *Bp = sB + kr * BS + col;
#pragma unroll
      for (int j = 0; j < L::MTW; ++j) {
        const int mt = wave + 4 * j;
        if (mt >= NPT) break;
        if (g == 0) {
          // the registers hold the wave's first tile, except after a wave's second tile
          if (j > 0) load_lift(mt);
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
              for (int b = 0; b < 3; ++b) sR[col * RG + b * Np + n] += acc[b][r];
            }
          }
          if (j > 0) load_lift(wave);   // the stress lift starts from the first tile
        } else if (g == 1) {
          if (j > 0) load_lift(mt);
          f4 acc[4][3];
#pragma unroll
          for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
              acc[f][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
              for (int kb = 0; kb < KBL; ++kb)
#pragma unroll
                for (int sb = 0; sb < 4; ++sb) {
                  const int kt = 4 * kb + sb;
                  if (kt < KTL)
                    acc[f][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aL[f][kb][sb], Bp[(b * KB + f * Nfp + 4 * kt) * BS],
                                                                     acc[f][b], 0, 0, 0);
                }
            }
          float nx[4], ny[4], nz[4];
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            nx[f] = sFg[L::FGS * col + 4 * f + 1];
            ny[f] = sFg[L::FGS * col + 4 * f + 2];
            nz[f] = sFg[L::FGS * col + 4 * f + 3];
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float xx = 0.f, yy = 0.f, zz = 0.f, yz = 0.f, xz = 0.f, xy = 0.f;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
              const float px = acc[f][0][r], py = acc[f][1][r], pz = acc[f][2][r];
              xx += nx[f] * px;
              yy += ny[f] * py;
              zz += nz[f] * pz;
              yz += ny[f] * pz + nz[f] * py;
              xz += nx[f] * pz + nz[f] * px;
              xy += nx[f] * py + ny[f] * px;
            }
            shr[j][0][r] = yz;
            shr[j][1][r] = xz;
            shr[j][2][r] = xy;
            const int n = mt * 16 + 4 * kr + r;
            if (n < Np) {
              float *rr = sR + col * RG + n;
              rr[0] += xx;
              rr[Np] += yy;
              rr[2 * Np] += zz;
            }
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int n = mt * 16 + 4 * kr + r;
            if (n < Np) {
#pragma unroll
              for (int b = 0; b < 3; ++b) sR[col * RG + b * Np + n] += shr[j][b][r];
            }
          }
        }
      }
    }
    __syncthreads();
    if constexpr (CAP) {
      const brsrc rY = mkrsrc(yout + (size_t)k0 * RowF, tbytes);
      for (int i = tid; i < TB * 3 * Np; i += 256) {
        const int e = i / (3 * Np), r = i - e * (3 * Np);
        bst(rY, sR[e * RG + r], (unsigned)(e * RowF + 3 * g * Np + r) * 4u);
      }
    }

    // ---- 6: S_b = map_g(Vq rhs_{3g+b}) (M = cubature points, N = elements, K = nodes)
    {
      const float *Bp = sR + col * RG + kr;
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

    // ---- 7: rhs_{3g+b} = Pq S_b (M = nodes, K = cubature points)
    for (int mt = wave; mt < NPT; mt += 4) {
      if (mt != wave) load_pq(mt);
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
          for (int b = 0; b < 3; ++b) sR[col * RG + b * Np + n] = acc[b][r];
        }
      }
      if (mt != wave && g < 2) load_pq(mt0);   // the next group starts from the first tile
    }
    __syncthreads();

    // ---- 8: LSRK for the group's rows: res = fa res + fdt (rhs + source); Q_out = Q_in + fb res
    elas_lsrk<N, RG>(g, in, sR, sQ, sSlot, src, has_src, ft, fa, fb, fdt, store_res, rRes, rOut, tid);
    // the next group's phase 3 writes F and the flux region, which nothing reads now; its
    // phase 4 writes rhs_g after the barrier that ends phase 3
  }
}


// The folded form (the default forward stage).  The velocities take one pass (A, B, C, D of
// the header).  The stresses take one A and B for both groups: the E F parts of the six stresses,
// and per face f and component c the product P = [Vq LIFT_f] w_c, which enters three stresses
// with the face's normal (sxx += nx P_x, sxz += nz P_x, sxy += ny P_x, ...), so w is lifted once
// for both groups; then C and D for the normal and for the shear stresses.
template <int N>
__global__ __launch_bounds__(256, N <= 4 ? 3 : N == 5 ? 2 : 1) void k_fused_elas_cq(FusedMesh m, FusedOps o, ElasMedia md,
                                                                                  const float *Qin, float *Qout, float *res,
                                                                                  SrcArgs src, float fa, float fb, float fdt,
                                                                                  float ft, const float *ftime_dev, int store_res) {
  using L = ElasCqLayout<N>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, Nq = L::Nq, TB = L::TB, RowF = L::RowF;
  constexpr int NQT = L::NQT, NQP = L::NQP, NPT = L::NPT, Np1 = L::Np1, QTW = L::QTW, MTW = L::MTW;
  constexpr int KT1 = L::KT1, KTL = L::KTL, KTC = L::KTC, KBC = L::KBC, KTP = L::KTP, KBP = L::KBP;
  constexpr int R1 = L::R1, RT = L::RT, RBV = L::RBV, BS = L::BS, RG = L::RG;
  extern __shared__ float lds[];
  float *sQ = lds + L::OQ;                 // [TB][RowF]
  float *sG = lds + L::OG;                 // vgeo [TB][9]
  float *sFg = sG + TB * 9;                // fgeo [TB][FGS]
  float *sW = lds + L::OW;                 // B operands (A, B), S [3][TB][SQS] (B, C)
  float *sR = lds + L::OR;                 // [TB][RG] (C, D)
  int *sSlot = (int *)(lds + L::OSL);
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int col = lane & 15, kr = lane >> 4;
  const int k0 = tile_index(m.xcd_chunk) * TB, nk = min(TB, m.K - k0);
  const unsigned tbytes = (unsigned)nk * RowF * 4u;
  if (ftime_dev) ft = *ftime_dev;
  static_assert(TB == 16, "the per-tile source table is built for 16-element tiles");
  const int2 stile = src_tile(src, k0, nk);
  const bool has_src = stile.y != 0;

  // ---- 1: loads, then the LDS stores
  ElasTileIn<N> in;
  elas_tile_load<N>(in, m, o, md, Qin, res, fa, k0, nk, tid);
  const brsrc rRes = mkrsrc(res + (size_t)k0 * RowF, tbytes);
  const brsrc rOut = mkrsrc(Qout + (size_t)k0 * RowF, tbytes);
  elas_tile_store<N>(in, sQ, sG, sFg, tid);
  if (has_src && tid < TB) sSlot[tid] = src_slot_tile(src, stile, k0, tid);
  __syncthreads();

  // ---- 2: penalty flux at the face nodes
  elas_flux<N>(in, sQ, sFg, md, k0, tid);

  // MFMA A operands.  The folded operator's blocks of the wave's cubature tile stay in registers
  // for both passes when the wave has one tile (N <= 4), and the Pq blocks of its node tile when
  // it has one (N <= 5); otherwise they are loaded per tile.  The loads are unconditional
  // (clamped to the last tile): loads under a wave test leave the compiler unsure how many are in
  // flight where the paths join.
  const brsrc rAC = mkrsrc(o.VqCBlk, (unsigned)NQT * KBC * 1024u);
  const brsrc rAP = mkrsrc(o.PqBlk, (unsigned)NPT * KBP * 1024u);
  constexpr bool KEEPC = QTW == 1, KEEPP = MTW == 1;
  f4 aC[KBC], aP[KBP];
  auto load_c = [&](int qt) {
    qt = qt < NQT ? qt : NQT - 1;
#pragma unroll
    for (int kb = 0; kb < KBC; ++kb) aC[kb] = bld_blk<KTC>(rAC, lane * 16u, (unsigned)(qt * KBC + kb) * 1024u, kb);
  };
  auto load_p = [&](int mt) {
    mt = mt < NPT ? mt : NPT - 1;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<KTP>(rAP, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
  };
  if constexpr (KEEPC) load_c(wave);
  if constexpr (KEEPP) load_p(wave);

  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf, f3 = m3 / Nfp;
  const bool on3 = g3 < L::NG3;
  const int g1 = tid / L::GS1, n1 = tid - g1 * L::GS1;
  const bool on1 = n1 < Np1;
  const float *Bp = sW + kr * BS + col;     // the lane's B operand reads (row k, element col)

  // one MFMA K-step of the folded operator: acc += A[kt] x (B rows rb + 4 (kt - k0) .. + 3)
  auto kstep = [&](f4 &acc, int kt, int rb) {
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(aC[kt / 4][kt % 4], Bp[rb * BS], acc, 0, 0, 0);
  };
  // S_b = map(acc) -> LDS for the wave's cubature tiles (points q >= Nq are zero)
  constexpr int SQS = L::SQS;
  auto store_s = [&](const f4 (&y)[QTW][3], auto &&map) {
#pragma unroll
    for (int t = 0; t < QTW; ++t) {
      const int qt = wave + 4 * t;
      if (qt >= NQT) break;
      f4 sv[3];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        float s3[3];
        map(t, r, y[t][0][r], y[t][1][r], y[t][2][r], s3);
        const bool qin = q < Nq;
#pragma unroll
        for (int b = 0; b < 3; ++b) sv[b][r] = qin ? s3[b] : 0.f;
      }
#pragma unroll
      for (int b = 0; b < 3; ++b) *(f4 *)(sW + (b * TB + col) * SQS + qt * 16 + 4 * kr) = sv[b];
    }
  };
  // C: rhs_b = Pq S_b (M = nodes, K = cubature points) -> sR
  auto pq = [&]() {
#pragma unroll
    for (int j = 0; j < MTW; ++j) {
      const int mt = wave + 4 * j;
      if (mt >= NPT) break;
      if constexpr (!KEEPP) load_p(mt);
      f4 pc[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) pc[b] = f4{0.f, 0.f, 0.f, 0.f};
      const float *Sp = sW + col * SQS + kr;
#pragma unroll
      for (int kb = 0; kb < KBP; ++kb)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < KTP) {
#pragma unroll
            for (int b = 0; b < 3; ++b)
              pc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[kb][sb], Sp[b * TB * SQS + kt * 4], pc[b], 0, 0, 0);
          }
        }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int b = 0; b < 3; ++b) sR[col * RG + b * Np + n] = pc[b][r];
        }
      }
    }
  };
  auto lsrk = [&](int g) {
    elas_lsrk<N, RG>(g, in, sR, sQ, sSlot, src, has_src, ft, fa, fb, fdt, store_res, rRes, rOut, tid);
  };

  // ================= velocities =================
  // ---- A: F (div sigma) -> rows b RBV + [0, Np1); velocity fluxes z -> rows b RBV + R1 + m
  if (on1) {
#pragma unroll
    for (int i = 0; i < L::EPT1; ++i) {
      const int e = g1 + L::NG1 * i;
      float Fv[6];
      elas_vol1<N>(0, sQ + e * RowF, sG + 9 * e, in.Sr, Fv);
#pragma unroll
      for (int b = 0; b < 3; ++b) sW[(b * RBV + n1) * BS + e] = Fv[b];
    }
  }
  if (on3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) {
#pragma unroll
        for (int b = 0; b < 3; ++b) sW[(b * RBV + R1 + m3) * BS + e] = in.fl[i][b];
      }
    }
  }
  {
    constexpr int NZ = L::NZV, PER = NZ / 3;
    for (int z = tid; z < NZ; z += 256) {
      const int b = z / PER, r = (z - b * PER) / BS, cc = z % BS;
      const int row = r < L::NZF ? Np1 + r : R1 + NfpNf + (r - L::NZF);
      sW[(b * RBV + row) * BS + cc] = 0.f;
    }
  }
  __syncthreads();
  // ---- B: Y_b = [Vq E | Vq LIFT_0 .. 3] [F_b; z_b]
  {
    f4 acc[QTW][3];
#pragma unroll
    for (int t = 0; t < QTW; ++t) {
      const int qt = wave + 4 * t;
#pragma unroll
      for (int b = 0; b < 3; ++b) acc[t][b] = f4{0.f, 0.f, 0.f, 0.f};
      if (qt < NQT) {
        if (!KEEPC) load_c(qt);
#pragma unroll
        for (int kt = 0; kt < KTC; ++kt) {
          const int row = kt < KT1 ? 4 * kt : R1 + ((kt - KT1) / KTL) * Nfp + 4 * ((kt - KT1) % KTL);
#pragma unroll
          for (int b = 0; b < 3; ++b) kstep(acc[t][b], kt, b * RBV + row);
        }
      }
    }
    __syncthreads();
    store_s(acc, [&](int t, int r, float y0, float y1, float y2, float (&s3)[3]) {
      s3[0] = in.rho[t][r] * y0;
      s3[1] = in.rho[t][r] * y1;
      s3[2] = in.rho[t][r] * y2;
    });
  }
  __syncthreads();
  // ---- C, D
  pq();
  __syncthreads();
  lsrk(0);

  // ================= stresses =================
  // ---- A: F of the normal (b = 0..2) and shear (b = 3..5) strain rates -> rows b R1 + [0, Np1);
  // the components of w -> rows 6 R1 + c RT + m.  (S of the velocities, which these words held,
  // was read before the last barrier; D reads sR and the Q_in tile only.)
  if (on1) {
#pragma unroll
    for (int i = 0; i < L::EPT1; ++i) {
      const int e = g1 + L::NG1 * i;
      float Fv[6];
      elas_vol1<N>(1, sQ + e * RowF, sG + 9 * e, in.Sr, Fv);
#pragma unroll
      for (int b = 0; b < 6; ++b) sW[(b * R1 + n1) * BS + e] = Fv[b];
    }
  }
  if (on3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) {
#pragma unroll
        for (int c = 0; c < 3; ++c) sW[(6 * R1 + c * RT + m3) * BS + e] = in.fl[i][3 + c];
      }
    }
  }
  {
    constexpr int NZ = L::NZT, NF = 6 * L::NZF * BS;
    for (int z = tid; z < NZ; z += 256) {
      const int cc = z % BS;
      int row = 0;
      if constexpr (L::NZF > 0) {
        if (z < NF) {
          const int b = z / (L::NZF * BS), r = (z - b * (L::NZF * BS)) / BS;
          row = b * R1 + Np1 + r;
        }
      }
      if (z >= NF) {
        const int zz = z - NF, c = zz / (4 * BS), r = (zz - c * (4 * BS)) / BS;
        row = 6 * R1 + c * RT + NfpNf + r;
      }
      sW[row * BS + cc] = 0.f;
    }
  }
  __syncthreads();
  // ---- B: Y_n, Y_s (normal, shear) = [Vq E] F + sum_f normal combinations of [Vq LIFT_f] w_c
  f4 yn[QTW][3], ys[QTW][3];
  {
    float nx[4], ny[4], nz[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      nx[f] = sFg[L::FGS * col + 4 * f + 1];
      ny[f] = sFg[L::FGS * col + 4 * f + 2];
      nz[f] = sFg[L::FGS * col + 4 * f + 3];
    }
#pragma unroll
    for (int t = 0; t < QTW; ++t) {
      const int qt = wave + 4 * t;
#pragma unroll
      for (int b = 0; b < 3; ++b) yn[t][b] = ys[t][b] = f4{0.f, 0.f, 0.f, 0.f};
      if (qt < NQT) {
        if (!KEEPC) load_c(qt);
#pragma unroll
        for (int kt = 0; kt < KT1; ++kt) {
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            kstep(yn[t][b], kt, b * R1 + 4 * kt);
            kstep(ys[t][b], kt, (3 + b) * R1 + 4 * kt);
          }
        }
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            f4 p = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < KTL; ++kk) kstep(p, KT1 + f * KTL + kk, 6 * R1 + c * RT + f * Nfp + 4 * kk);
            // c = x: sxx nx, sxz nz, sxy ny; c = y: syy ny, syz nz, sxy nx; c = z: szz nz, syz ny, sxz nx
            const float nn = c == 0 ? nx[f] : c == 1 ? ny[f] : nz[f];
            const float na = c == 0 ? nz[f] : c == 1 ? nz[f] : ny[f];
            const float nb = c == 0 ? ny[f] : c == 1 ? nx[f] : nx[f];
            const int sa = c == 0 ? 1 : 0, sb2 = c == 2 ? 1 : 2;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              yn[t][c][r] += nn * p[r];
              ys[t][sa][r] += na * p[r];
              ys[t][sb2][r] += nb * p[r];
            }
          }
      }
    }
  }
  __syncthreads();
  // ---- normal stresses: S = the isotropic stiffness of Y_n; C; then S of the shear stresses
  // (mu Y_s) and D of the normal ones between the same barriers
  store_s(yn, [&](int t, int r, float y0, float y1, float y2, float (&s3)[3]) {
    const float A = 2.f * in.mu[t][r] + in.lam[t][r], l = in.lam[t][r];
    s3[0] = A * y0 + l * y1 + l * y2;
    s3[1] = A * y1 + l * y0 + l * y2;
    s3[2] = A * y2 + l * y0 + l * y1;
  });
  __syncthreads();
  pq();
  __syncthreads();
  store_s(ys, [&](int t, int r, float y0, float y1, float y2, float (&s3)[3]) {
    s3[0] = in.mu[t][r] * y0;
    s3[1] = in.mu[t][r] * y1;
    s3[2] = in.mu[t][r] * y2;
  });
  lsrk(1);
  __syncthreads();
  pq();
  __syncthreads();
  lsrk(2);
}

template <int N, bool CAP>
static hipError_t fused_elas_n(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb, float fdt,
                               const float *ftd, float ft, const SrcArgs &s, int store_res, float *yout) {
  using L = ElasFusedLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if ((unsigned long long)c->K * 36 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused_elas<N, CAP>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  FusedMesh m{c->K, c->vgeo, c->fgeo, c->vmapP, c->xcd_chunk};
  ElasMedia md{c->rhoinv, c->lambda, c->mu, c->tau_v, c->tau_s};
  hipLaunchKernelGGL((k_fused_elas<N, CAP>), dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream, m, fused_ops(c), md,
                     Qin, Qout, c->res, s, fa, fb, fdt, ft, ftd, store_res, yout);
  return hipGetLastError();
}

template <int N>
static hipError_t fused_elas_cq_n(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb, float fdt,
                                  const float *ftd, float ft, const SrcArgs &s, int store_res) {
  using L = ElasCqLayout<N>;
  const size_t sh = sizeof(float) * L::floats;
  if ((unsigned long long)c->K * 36 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused_elas_cq<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  FusedMesh m{c->K, c->vgeo, c->fgeo, c->vmapP, c->xcd_chunk};
  ElasMedia md{c->rhoinv, c->lambda, c->mu, c->tau_v, c->tau_s};
  hipLaunchKernelGGL((k_fused_elas_cq<N>), dim3((c->K + L::TB - 1) / L::TB), dim3(256), sh, c->stream, m, fused_ops(c), md,
                     Qin, Qout, c->res, s, fa, fb, fdt, ft, ftd, store_res);
  return hipGetLastError();
}

template <int N>
static hipError_t fused_elas_any(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb, float fdt,
                                 const float *ftd, float ft, const SrcArgs &s, int store_res, float *yout) {
  if (yout) return fused_elas_n<N, true>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, yout);
  if (elas_form(c) == BBW_ELAS_LIFT_CUBATURE) return fused_elas_cq_n<N>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res);
  return fused_elas_n<N, false>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, nullptr);
}

hipError_t launch_fused_elas(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb, float fdt,
                             const float *ftd, float ft, int store_res, float *yout) {
  SrcArgs s = ctx_src(c);
  switch (c->N) {
    case 1: return fused_elas_any<1>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, yout);
    case 2: return fused_elas_any<2>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, yout);
    case 3: return fused_elas_any<3>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, yout);
    case 4: return fused_elas_any<4>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, yout);
    case 5: return fused_elas_any<5>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, yout);
    case 6: return fused_elas_any<6>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, store_res, yout);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
