/*
 * CPU restatement of the BB-WADG hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: tests, the smoke check and the CPU-baseline leg of
 * the benchmark call it to check (and time beside) the HIP library.  It is
 * never linked into the product.  It restates, loop for loop, the reference's
 * OKL kernels in plain C, once in FP64 (the host oracle) and once in FP32
 * (same arithmetic type as the reference's dfloat = float), with OpenMP over
 * elements in place of the OKL outer0 loop:
 *
 *   orc_volume_*      rk_volume_bern      cpp/okl/WaveKernels.okl:344-474
 *   orc_surface_*     rk_surface_bern     cpp/okl/WaveKernels.okl:479-647
 *   orc_update_wadg_* rk_update_WADG      cpp/okl/WaveKernelsWADG.okl:1760-1887
 *   orc_update_*      rk_update           cpp/okl/WaveKernels.okl:262-313
 *   orc_volume_elas_*  BB restatement of rk_volume_elas  elas_cpp/okl/ElasKernelsWADG.okl:25-124
 *   orc_surface_elas_* BB-lift restatement of rk_surface_elas  elas_cpp/okl/ElasKernelsWADG.okl:128-284
 *   orc_update_elas_*  rk_update_elas     elas_cpp/okl/ElasKernelsWADG.okl:286-433
 *   orc_rk_*          Wave_RK + RK_step_WADG_subelem  cpp/src/WaveOKL3d.cpp:2014-2109
 *
 * Layouts are the reference's: Q/rhs/res [K][Nf][Np]; vgeo [K][9]; fgeo
 * [K][4][4]; vmapP [K][4*Nfp] (field-strided flat index into Q); Fmask [4][Nfp];
 * D_ids [4 j][Np][4 m]; Dvals [Np][4]; EEL/L0 ids and values column-major
 * (entry (i,j) at i + j*rows); Vq column-major Nq x Np; Pq column-major Np x Nq;
 * c2q, Jq [K][Nq].
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -shared).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NFACES 4

#define DEFINE_ORACLE(R, SFX)                                                            \
                                                                                         \
void orc_volume_##SFX(int K, int Np, const R *vgeo, const int *Dids, const R *Dvals,     \
                      const R *Q, R *rhsQ)                                               \
{                                                                                        \
  const int Nf = 4;                                                                      \
  _Pragma("omp parallel for schedule(static) if(K >= 2048)")                                           \
  for (int k = 0; k < K; ++k) {                                                          \
    R sp[512], sUr[512], sUs[512], sUt[512];                                             \
    const R *G = vgeo + 9 * k;                                                           \
    const R rx = G[0], ry = G[1], rz = G[2], sx = G[3], sy = G[4], sz = G[5];            \
    const R tx = G[6], ty = G[7], tz = G[8];                                             \
    for (int n = 0; n < Np; ++n) {                                                       \
      const int id = n + k * Np * Nf;                                                    \
      sp[n] = Q[id];                                                                     \
      const R un = Q[id + Np], vn = Q[id + 2 * Np], wn = Q[id + 3 * Np];                 \
      sUr[n] = un * rx + vn * ry + wn * rz;                                              \
      sUs[n] = un * sx + vn * sy + wn * sz;                                              \
      sUt[n] = un * tx + vn * ty + wn * tz;                                              \
    }                                                                                    \
    for (int n = 0; n < Np; ++n) {                                                       \
      const R *Dv = Dvals + 4 * n;                                                       \
      const int *D1i = Dids + 4 * n;                                                     \
      const int *D2i = Dids + 4 * (Np + n);                                              \
      const int *D3i = Dids + 4 * (2 * Np + n);                                          \
      const int *D4i = Dids + 4 * (3 * Np + n);                                          \
      const R p1 = Dv[0]*sp[D1i[0]] + Dv[1]*sp[D2i[0]] + Dv[2]*sp[D3i[0]] + Dv[3]*sp[D4i[0]]; \
      const R p2 = Dv[0]*sp[D1i[1]] + Dv[1]*sp[D2i[1]] + Dv[2]*sp[D3i[1]] + Dv[3]*sp[D4i[1]]; \
      const R p3 = Dv[0]*sp[D1i[2]] + Dv[1]*sp[D2i[2]] + Dv[2]*sp[D3i[2]] + Dv[3]*sp[D4i[2]]; \
      const R p4 = Dv[0]*sp[D1i[3]] + Dv[1]*sp[D2i[3]] + Dv[2]*sp[D3i[3]] + Dv[3]*sp[D4i[3]]; \
      R dU1 = Dv[0] * (sUr[D1i[1]] + sUs[D1i[2]] + sUt[D1i[3]]);                         \
      dU1 += Dv[1] * (sUr[D2i[1]] + sUs[D2i[2]] + sUt[D2i[3]]);                          \
      dU1 += Dv[2] * (sUr[D3i[1]] + sUs[D3i[2]] + sUt[D3i[3]]);                          \
      dU1 += Dv[3] * (sUr[D4i[1]] + sUs[D4i[2]] + sUt[D4i[3]]);                          \
      R dU2 = Dv[0] * (sUr[D1i[0]] + sUs[D1i[0]] + sUt[D1i[0]]);                         \
      dU2 += Dv[1] * (sUr[D2i[0]] + sUs[D2i[0]] + sUt[D2i[0]]);                          \
      dU2 += Dv[2] * (sUr[D3i[0]] + sUs[D3i[0]] + sUt[D3i[0]]);                          \
      dU2 += Dv[3] * (sUr[D4i[0]] + sUs[D4i[0]] + sUt[D4i[0]]);                          \
      const R dpdr = (R)0.5 * (p2 - p1), dpds = (R)0.5 * (p3 - p1);                      \
      const R dpdt = (R)0.5 * (p4 - p1), divU = (R)0.5 * (dU1 - dU2);                    \
      const R dpdx = rx * dpdr + sx * dpds + tx * dpdt;                                  \
      const R dpdy = ry * dpdr + sy * dpds + ty * dpdt;                                  \
      const R dpdz = rz * dpdr + sz * dpds + tz * dpdt;                                  \
      const int id = n + k * Nf * Np;                                                    \
      rhsQ[id] = -divU;                                                                  \
      rhsQ[id + Np] = -dpdx;                                                             \
      rhsQ[id + 2 * Np] = -dpdy;                                                         \
      rhsQ[id + 3 * Np] = -dpdz;                                                         \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
void orc_surface_##SFX(int K, int Np, int Nfp, const R *fgeo, const int *Fmask,          \
                       const int *vmapP, const int *EEL_ids, const R *EEL_vals,          \
                       int EEL_nnz, const int *L0_ids, const R *L0_vals, int L0_nnz,     \
                       const R *Q, R *rhsQ)                                              \
{                                                                                        \
  const int Nf = 4, NfpNf = NFACES * Nfp;                                                \
  _Pragma("omp parallel for schedule(static) if(K >= 2048)")                                           \
  for (int k = 0; k < K; ++k) {                                                          \
    R s_pflux[256], s_Uflux[256], s_ptmp[256], s_Utmp[256], s_nxyz[16];                  \
    for (int f = 0; f < NFACES; ++f) {                                                   \
      const R *fg = fgeo + 4 * (f + NFACES * k);                                         \
      s_nxyz[4 * f + 0] = fg[1]; s_nxyz[4 * f + 1] = fg[2];                              \
      s_nxyz[4 * f + 2] = fg[3]; s_nxyz[4 * f + 3] = fg[0];                              \
    }                                                                                    \
    for (int n = 0; n < NfpNf; ++n) {                                                    \
      const int f = n / Nfp;                                                             \
      int idM = Fmask[n] + k * Np * Nf;                                                  \
      int idP = vmapP[n + k * NfpNf];                                                    \
      const int isBoundary = (idM == idP);                                               \
      const R pM = Q[idM], uM = Q[idM + Np], vM = Q[idM + 2 * Np], wM = Q[idM + 3 * Np]; \
      const R pP = Q[idP], uP = Q[idP + Np], vP = Q[idP + 2 * Np], wP = Q[idP + 3 * Np]; \
      R pjump = pP - pM;                                                                 \
      R Unjump = (uP - uM) * s_nxyz[4 * f] + (vP - vM) * s_nxyz[4 * f + 1]               \
                 + (wP - wM) * s_nxyz[4 * f + 2];                                        \
      if (isBoundary) { pjump = (R)-2 * pM; Unjump = 0; }                                \
      const R Fscale = s_nxyz[4 * f + 3];                                                \
      s_pflux[n] = (R)0.5 * (pjump - Unjump) * Fscale;                                   \
      s_Uflux[n] = (R)0.5 * (Unjump - pjump) * Fscale;                                   \
    }                                                                                    \
    for (int n = 0; n < NfpNf; ++n) {                                                    \
      const int f = n / Nfp, nt = n % Nfp;                                               \
      R val1 = 0, val2 = 0;                                                              \
      for (int j = 0; j < L0_nnz; ++j) {                                                 \
        const R L0j = L0_vals[nt + j * Nfp];                                             \
        const int id = L0_ids[nt + j * Nfp] + f * Nfp;                                   \
        val1 += L0j * s_pflux[id];                                                       \
        val2 += L0j * s_Uflux[id];                                                       \
      }                                                                                  \
      s_ptmp[n] = val1; s_Utmp[n] = val2;                                                \
    }                                                                                    \
    for (int n = 0; n < Np; ++n) {                                                       \
      const int id = n + k * Np * Nf;                                                    \
      R val1 = rhsQ[id], val2 = rhsQ[id + Np], val3 = rhsQ[id + 2 * Np];                 \
      R val4 = rhsQ[id + 3 * Np];                                                        \
      for (int j = 0; j < EEL_nnz; ++j) {                                                \
        const int col = EEL_ids[n + j * Np];                                             \
        const R Ev = EEL_vals[n + j * Np];                                               \
        const int fc = col / Nfp;                                                        \
        const R Uf = s_Utmp[col];                                                        \
        val1 += Ev * s_ptmp[col];                                                        \
        val2 += Ev * Uf * s_nxyz[4 * fc];                                                \
        val3 += Ev * Uf * s_nxyz[4 * fc + 1];                                            \
        val4 += Ev * Uf * s_nxyz[4 * fc + 2];                                            \
      }                                                                                  \
      rhsQ[id] = val1; rhsQ[id + Np] = val2; rhsQ[id + 2 * Np] = val3;                   \
      rhsQ[id + 3 * Np] = val4;                                                          \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
/* fields in [0, nproj) are projected through Pq diag(w) Vq; fields >= nproj   */       \
/* (velocity, when the projection is skipped) pass through unchanged.  With     */       \
/* nproj = 4 this is rk_update_WADG exactly.  src (may be NULL): per-element    */       \
/* BB coefficients added to field 0 as ftime*src after the projection.          */       \
void orc_update_wadg_##SFX(int K, int Np, int Nq, int nproj, const R *Vq, const R *Pq,   \
                           const R *Jq, const R *c2q, R fa, R fb, R fdt, R ftime,        \
                           const R *src, const R *rhsQ, R *resQ, R *Q)                   \
{                                                                                        \
  const int Nf = 4;                                                                      \
  _Pragma("omp parallel for schedule(static) if(K >= 2048)")                                           \
  for (int k = 0; k < K; ++k) {                                                          \
    R sq[4][1024], r[4][256];                                                            \
    for (int q = 0; q < Nq; ++q) {                                                       \
      R acc[4] = {0, 0, 0, 0};                                                           \
      for (int j = 0; j < Np; ++j) {                                                     \
        const R v = Vq[q + j * Nq];                                                      \
        for (int f = 0; f < nproj; ++f) acc[f] += v * rhsQ[j + f * Np + k * Np * Nf];    \
      }                                                                                  \
      const R c2 = c2q[q + k * Nq];                                                      \
      const R invJ = Jq ? (R)1 / Jq[q + k * Nq] : (R)1;                                  \
      sq[0][q] = acc[0] * invJ * c2;                                                     \
      for (int f = 1; f < nproj; ++f) sq[f][q] = acc[f] * invJ;                          \
    }                                                                                    \
    for (int n = 0; n < Np; ++n) {                                                       \
      R acc[4] = {0, 0, 0, 0};                                                           \
      for (int j = 0; j < Nq; ++j) {                                                     \
        const R p = Pq[n + j * Np];                                                      \
        for (int f = 0; f < nproj; ++f) acc[f] += p * sq[f][j];                          \
      }                                                                                  \
      for (int f = 0; f < nproj; ++f) r[f][n] = acc[f];                                  \
      for (int f = nproj; f < Nf; ++f) r[f][n] = rhsQ[n + f * Np + k * Np * Nf];         \
      if (src) r[0][n] += ftime * src[n + k * Np];                                       \
    }                                                                                    \
    for (int n = 0; n < Np; ++n) {                                                       \
      for (int f = 0; f < Nf; ++f) {                                                     \
        const int id = n + f * Np + k * Np * Nf;                                         \
        R res = resQ[id];                                                                \
        res = fa * res + fdt * r[f][n];                                                  \
        resQ[id] = res;                                                                  \
        Q[id] += fb * res;                                                               \
      }                                                                                  \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
void orc_update_##SFX(long n, R fa, R fb, R fdt, const R *rhsQ, R *resQ, R *Q)           \
{                                                                                        \
  _Pragma("omp parallel for schedule(static) if(n >= 65536)")                                           \
  for (long i = 0; i < n; ++i) {                                                         \
    R res = resQ[i];                                                                     \
    res = fa * res + fdt * rhsQ[i];                                                      \
    resQ[i] = res;                                                                       \
    Q[i] += fb * res;                                                                    \
  }                                                                                      \
}

DEFINE_ORACLE(double, f64)
DEFINE_ORACLE(float, f32)

/* --------------------------------------------------------------------------
 * Elastic velocity-stress, 9 fields (v1 v2 v3 sxx syy szz syz sxz sxy).
 *
 * Volume: the reference's rk_volume_elas (ElasKernelsWADG.okl:25-124) applies
 * nodal Dr/Ds/Dt; in the BB basis the same derivatives come from the 4-nnz
 * barycentric operators: d/dr = (D2 - D1)/2, d/ds = (D3 - D1)/2,
 * d/dt = (D4 - D1)/2 (lam2 = (1+r)/2, lam1 = -(1+r+s+t)/2, ...).
 * Surface: central flux with penalties tau_v, tau_s exactly as
 * rk_surface_elas (ElasKernelsWADG.okl:163-240), boundary Q+ = 0, then the BB
 * lift EEL . blkdiag(L0) in place of the dense nodal LIFT.
 * ------------------------------------------------------------------------ */
#define DEFINE_ELAS(R, SFX)                                                              \
void orc_volume_elas_##SFX(int K, int Np, const R *vgeo, const int *Dids,                \
                           const R *Dvals, const R *Q, R *rhsQ)                          \
{                                                                                        \
  const int Nf = 9;                                                                      \
  _Pragma("omp parallel for schedule(static) if(K >= 2048)")                                           \
  for (int k = 0; k < K; ++k) {                                                          \
    const R *G = vgeo + 9 * k;                                                           \
    const R *Qk = Q + (long)k * Nf * Np;                                                 \
    for (int n = 0; n < Np; ++n) {                                                       \
      const R *Dv = Dvals + 4 * n;                                                       \
      R Qx[9], Qy[9], Qz[9];                                                             \
      for (int fl = 0; fl < Nf; ++fl) {                                                  \
        const R *s = Qk + fl * Np;                                                       \
        R pm[4];                                                                         \
        for (int m = 0; m < 4; ++m) {                                                    \
          pm[m] = Dv[0] * s[Dids[4 * n + m]] + Dv[1] * s[Dids[4 * (Np + n) + m]]         \
                + Dv[2] * s[Dids[4 * (2 * Np + n) + m]] + Dv[3] * s[Dids[4 * (3 * Np + n) + m]]; \
        }                                                                                \
        const R dr = (R)0.5 * (pm[1] - pm[0]), ds = (R)0.5 * (pm[2] - pm[0]);            \
        const R dt = (R)0.5 * (pm[3] - pm[0]);                                           \
        Qx[fl] = G[0] * dr + G[3] * ds + G[6] * dt;                                      \
        Qy[fl] = G[1] * dr + G[4] * ds + G[7] * dt;                                      \
        Qz[fl] = G[2] * dr + G[5] * ds + G[8] * dt;                                      \
      }                                                                                  \
      R *o = rhsQ + (long)k * Nf * Np + n;                                               \
      o[0 * Np] = Qx[3] + Qy[8] + Qz[7];                                                 \
      o[1 * Np] = Qx[8] + Qy[4] + Qz[6];                                                 \
      o[2 * Np] = Qx[7] + Qy[6] + Qz[5];                                                 \
      o[3 * Np] = Qx[0];                                                                 \
      o[4 * Np] = Qy[1];                                                                 \
      o[5 * Np] = Qz[2];                                                                 \
      o[6 * Np] = Qy[2] + Qz[1];                                                         \
      o[7 * Np] = Qx[2] + Qz[0];                                                         \
      o[8 * Np] = Qx[1] + Qy[0];                                                         \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
void orc_surface_elas_##SFX(int K, int Np, int Nfp, const R *fgeo, const int *Fmask,     \
                            const int *vmapP, const int *EEL_ids, const R *EEL_vals,     \
                            int EEL_nnz, const int *L0_ids, const R *L0_vals,            \
                            int L0_nnz, R tau_v, R tau_s, const R *Q, R *rhsQ)           \
{                                                                                        \
  const int Nf = 9, NfpNf = NFACES * Nfp;                                                \
  _Pragma("omp parallel for schedule(static) if(K >= 2048)")                                           \
  for (int k = 0; k < K; ++k) {                                                          \
    R flux[9][256], tmp[9][256];                                                         \
    for (int i = 0; i < NfpNf; ++i) {                                                    \
      const int f = i / Nfp;                                                             \
      int idM = Fmask[i] + k * Np * Nf;                                                  \
      int idP = vmapP[i + k * NfpNf];                                                    \
      const int isB = (idM == idP);                                                      \
      const R *fg = fgeo + 4 * (f + NFACES * k);                                         \
      const R Fscale = fg[0], nx = fg[1], ny = fg[2], nz = fg[3];                        \
      R dQ[9];                                                                           \
      for (int fl = 0; fl < Nf; ++fl) {                                                  \
        dQ[fl] = -Q[idM];                                                                \
        if (!isB) dQ[fl] += Q[idP];                                                      \
        idM += Np; idP += Np;                                                            \
      }                                                                                  \
      R fc[9];                                                                           \
      fc[0] = dQ[3] * nx + dQ[8] * ny + dQ[7] * nz;                                      \
      fc[1] = dQ[8] * nx + dQ[4] * ny + dQ[6] * nz;                                      \
      fc[2] = dQ[7] * nx + dQ[6] * ny + dQ[5] * nz;                                      \
      fc[3] = dQ[0] * nx;                                                                \
      fc[4] = dQ[1] * ny;                                                                \
      fc[5] = dQ[2] * nz;                                                                \
      fc[6] = dQ[2] * ny + dQ[1] * nz;                                                   \
      fc[7] = dQ[2] * nx + dQ[0] * nz;                                                   \
      fc[8] = dQ[1] * nx + dQ[0] * ny;                                                   \
      R fl_[9];                                                                          \
      fl_[0] = fc[0] + tau_v * (fc[3] * nx + fc[8] * ny + fc[7] * nz);                   \
      fl_[1] = fc[1] + tau_v * (fc[8] * nx + fc[4] * ny + fc[6] * nz);                   \
      fl_[2] = fc[2] + tau_v * (fc[7] * nx + fc[6] * ny + fc[5] * nz);                   \
      fl_[3] = fc[3] + tau_s * (fc[0] * nx);                                             \
      fl_[4] = fc[4] + tau_s * (fc[1] * ny);                                             \
      fl_[5] = fc[5] + tau_s * (fc[2] * nz);                                             \
      fl_[6] = fc[6] + tau_s * (fc[2] * ny + fc[1] * nz);                                \
      fl_[7] = fc[7] + tau_s * (fc[2] * nx + fc[0] * nz);                                \
      fl_[8] = fc[8] + tau_s * (fc[1] * nx + fc[0] * ny);                                \
      for (int fl = 0; fl < Nf; ++fl) flux[fl][i] = fl_[fl] * (R)0.5 * Fscale;           \
    }                                                                                    \
    for (int i = 0; i < NfpNf; ++i) {                                                    \
      const int f = i / Nfp, nt = i % Nfp;                                               \
      for (int fl = 0; fl < Nf; ++fl) {                                                  \
        R v = 0;                                                                         \
        for (int j = 0; j < L0_nnz; ++j)                                                 \
          v += L0_vals[nt + j * Nfp] * flux[fl][L0_ids[nt + j * Nfp] + f * Nfp];         \
        tmp[fl][i] = v;                                                                  \
      }                                                                                  \
    }                                                                                    \
    for (int n = 0; n < Np; ++n) {                                                       \
      for (int fl = 0; fl < Nf; ++fl) {                                                  \
        R v = 0;                                                                         \
        for (int j = 0; j < EEL_nnz; ++j)                                                \
          v += EEL_vals[n + j * Np] * tmp[fl][EEL_ids[n + j * Np]];                      \
        rhsQ[n + fl * Np + (long)k * Nf * Np] += v;                                      \
      }                                                                                  \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
/* rk_update_elas (isotropic branch): velocity x rhoinv, stress = C : strain rate, */   \
/* then Pq, + ftime*srcw[field]*src (the reference puts it on field 0 only), LSRK. */   \
void orc_update_elas_##SFX(int K, int Np, int Nq, const R *Vq, const R *Pq,              \
                           const R *rhoinvq, const R *lambdaq, const R *muq,             \
                           R ftime, const R *src, const R *srcw, R fa, R fb, R fdt,      \
                           const R *rhsQ, R *resQ, R *Q)                                 \
{                                                                                        \
  const int Nf = 9;                                                                      \
  _Pragma("omp parallel for schedule(static) if(K >= 2048)")                                           \
  for (int k = 0; k < K; ++k) {                                                          \
    R sQ[9][1024];                                                                       \
    for (int i = 0; i < Nq; ++i) {                                                       \
      R rq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};                                             \
      for (int j = 0; j < Np; ++j) {                                                     \
        const R v = Vq[i + j * Nq];                                                      \
        for (int fl = 0; fl < Nf; ++fl) rq[fl] += v * rhsQ[j + fl * Np + (long)k * Nf * Np]; \
      }                                                                                  \
      const int id = i + k * Nq;                                                         \
      const R rho = rhoinvq[id], lambda = lambdaq[id], mu = muq[id];                     \
      sQ[0][i] = rho * rq[0]; sQ[1][i] = rho * rq[1]; sQ[2][i] = rho * rq[2];            \
      const R A = (R)2 * mu + lambda;                                                    \
      sQ[3][i] = A * rq[3] + lambda * rq[4] + lambda * rq[5];                            \
      sQ[4][i] = A * rq[4] + lambda * rq[3] + lambda * rq[5];                            \
      sQ[5][i] = A * rq[5] + lambda * rq[3] + lambda * rq[4];                            \
      sQ[6][i] = mu * rq[6]; sQ[7][i] = mu * rq[7]; sQ[8][i] = mu * rq[8];               \
    }                                                                                    \
    for (int i = 0; i < Np; ++i) {                                                       \
      R rQ[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};                                             \
      for (int j = 0; j < Nq; ++j) {                                                     \
        const R p = Pq[i + j * Np];                                                      \
        for (int fl = 0; fl < Nf; ++fl) rQ[fl] += p * sQ[fl][j];                         \
      }                                                                                  \
      if (src)                                                                           \
        for (int fl = 0; fl < Nf; ++fl) rQ[fl] += ftime * srcw[fl] * src[i + k * Np];    \
      for (int fl = 0; fl < Nf; ++fl) {                                                  \
        const long id = i + fl * Np + (long)k * Nf * Np;                                 \
        R res = resQ[id];                                                                \
        res = fa * res + fdt * rQ[fl];                                                   \
        resQ[id] = res;                                                                  \
        Q[id] += fb * res;                                                               \
      }                                                                                  \
    }                                                                                    \
  }                                                                                      \
}

DEFINE_ELAS(double, f64)
DEFINE_ELAS(float, f32)
