/*
 * CPU restatement of the curvilinear WADG kernels -- TEST INFRASTRUCTURE ONLY (the oracle; never
 * linked into the product).  Loop for loop, in FP64 and FP32, with OpenMP over elements:
 *
 *   orc_volume_skew_*     rk_volume_WADG_skew_combine   cpp/okl/WaveKernelsWADG.okl:514-682
 *   orc_surface_skew_*    rk_surface_WADG_skew_combine  cpp/okl/WaveKernelsWADG.okl:3024-3195
 *   orc_volume_planar_*   rk_volume_planar              cpp/okl/WaveKernelsWADG.okl:2203-2317
 *   orc_surface_planar_*  rk_surface_planar             cpp/okl/WaveKernelsWADG.okl:2320-2441
 *
 * The element lists (KlistCurved / KlistPlanar) select the elements; the nodal state and the
 * rhs are [K][4][Np].  Layouts are the reference's: vgeoq [K][9][Nq] (rx sx tx ry sy ty rz sz
 * tz), Jq [K][Nq], fgeoq [K][4 faces][4 (nx ny nz sJ)][Nfq], vgeo [K][9] (rx ry rz sx sy sz tx
 * ty tz), fgeo [K][4][4] (Fscale nx ny nz), J [K]; Vq/Vrq/Vsq/Vtq column-major Nq x Np; Pq/Prq/
 * Psq/Ptq column-major Np x Nq; VfqFace column-major Nfq x Nfp; Pfq column-major Np x 4Nfq;
 * Dr/Ds/Dt column-major Np x Np; LIFT column-major Np x 4Nfp.
 */
#include <math.h>

#define NFACES 4
#define MAXQ 512
#define MAXN 256

#define DEFINE_CURVED(R, SFX)                                                            \
void orc_volume_skew_##SFX(int Kl, const int *Klist, int Np, int Nq, const R *vgeoq,     \
                           const R *Jq, const R *Vq, const R *Vrq, const R *Vsq,          \
                           const R *Vtq, const R *Pq, const R *Prq, const R *Psq,         \
                           const R *Ptq, const R *Q, R *rhsQ)                             \
{                                                                                        \
  _Pragma("omp parallel for schedule(static) if(Kl >= 1024)")                             \
  for (int kk = 0; kk < Kl; ++kk) {                                                      \
    const int k = Klist[kk];                                                             \
    R spx[MAXQ], spy[MAXQ], spz[MAXQ], su[MAXQ], sv[MAXQ], sw[MAXQ];                      \
    const R *q = Q + (long)k * 4 * Np;                                                    \
    for (int n = 0; n < Nq; ++n) {                                                       \
      const R *g = vgeoq + (long)k * 9 * Nq + n;                                         \
      const R rx = g[0], sx = g[Nq], tx = g[2 * Nq], ry = g[3 * Nq], sy = g[4 * Nq];      \
      const R ty = g[5 * Nq], rz = g[6 * Nq], sz = g[7 * Nq], tz = g[8 * Nq];             \
      const R J = Jq[n + (long)k * Nq];                                                  \
      R dpdr = 0, dpds = 0, dpdt = 0, uq = 0, vq = 0, wq = 0;                            \
      for (int m = 0; m < Np; ++m) {                                                     \
        const R pm = q[m], um = q[Np + m], vm = q[2 * Np + m], wm = q[3 * Np + m];       \
        const R Vm = Vq[n + m * Nq];                                                     \
        uq += Vm * um; vq += Vm * vm; wq += Vm * wm;                                     \
        dpdr += Vrq[n + m * Nq] * pm;                                                    \
        dpds += Vsq[n + m * Nq] * pm;                                                    \
        dpdt += Vtq[n + m * Nq] * pm;                                                    \
      }                                                                                  \
      spx[n] = (rx * dpdr + sx * dpds + tx * dpdt) * J;                                  \
      spy[n] = (ry * dpdr + sy * dpds + ty * dpdt) * J;                                  \
      spz[n] = (rz * dpdr + sz * dpds + tz * dpdt) * J;                                  \
      su[n] = -(uq * rx + vq * ry + wq * rz) * J;                                        \
      sv[n] = -(uq * sx + vq * sy + wq * sz) * J;                                        \
      sw[n] = -(uq * tx + vq * ty + wq * tz) * J;                                        \
    }                                                                                    \
    R *o = rhsQ + (long)k * 4 * Np;                                                       \
    for (int n = 0; n < Np; ++n) {                                                       \
      R divU = 0, dpdx = 0, dpdy = 0, dpdz = 0;                                          \
      for (int m = 0; m < Nq; ++m) {                                                     \
        const R P = Pq[n + m * Np];                                                      \
        dpdx += P * spx[m]; dpdy += P * spy[m]; dpdz += P * spz[m];                      \
        divU += Prq[n + m * Np] * su[m];                                                 \
        divU += Psq[n + m * Np] * sv[m];                                                 \
        divU += Ptq[n + m * Np] * sw[m];                                                 \
      }                                                                                  \
      o[n] = -divU; o[Np + n] = -dpdx; o[2 * Np + n] = -dpdy; o[3 * Np + n] = -dpdz;     \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
void orc_surface_skew_##SFX(int Kl, const int *Klist, int Np, int Nfp, int Nfq,          \
                            const R *fgeoq, const R *VfqFace, const R *Pfq,               \
                            const int *Fmask, const int *vmapP, const R *Q, R *rhsQ)      \
{                                                                                        \
  _Pragma("omp parallel for schedule(static) if(Kl >= 1024)")                             \
  for (int kk = 0; kk < Kl; ++kk) {                                                      \
    const int k = Klist[kk];                                                             \
    R dp[NFACES][MAXN], du[NFACES][MAXN], dv[NFACES][MAXN], dw[NFACES][MAXN];            \
    R au[NFACES][MAXN], av[NFACES][MAXN], aw[NFACES][MAXN];                              \
    for (int f = 0; f < NFACES; ++f)                                                     \
      for (int n = 0; n < Nfp; ++n) {                                                    \
        const int m = n + f * Nfp;                                                       \
        int idM = Fmask[m] + k * Np * 4;                                                 \
        int idP = vmapP[m + (long)k * Nfp * NFACES];                                     \
        const int bnd = idM == idP;                                                      \
        const R pM = Q[idM], uM = Q[idM + Np], vM = Q[idM + 2 * Np], wM = Q[idM + 3 * Np]; \
        const R pP = Q[idP], uP = Q[idP + Np], vP = Q[idP + 2 * Np], wP = Q[idP + 3 * Np]; \
        dp[f][n] = bnd ? -2 * pM : pP - pM;                                              \
        du[f][n] = bnd ? 0 : uP - uM;                                                    \
        dv[f][n] = bnd ? 0 : vP - vM;                                                    \
        dw[f][n] = bnd ? 0 : wP - wM;                                                    \
        au[f][n] = bnd ? uM : (R)0.5 * (uP + uM);                                        \
        av[f][n] = bnd ? vM : (R)0.5 * (vP + vM);                                        \
        aw[f][n] = bnd ? wP : (R)0.5 * (wP + wM);                                        \
      }                                                                                  \
    R fp[NFACES][MAXN], fu[NFACES][MAXN], fv[NFACES][MAXN], fw[NFACES][MAXN];            \
    for (int f = 0; f < NFACES; ++f)                                                     \
      for (int n = 0; n < Nfq; ++n) {                                                    \
        R dpq = 0, duq = 0, dvq = 0, dwq = 0, auq = 0, avq = 0, awq = 0;                 \
        for (int j = 0; j < Nfp; ++j) {                                                  \
          const R V = VfqFace[n + j * Nfq];                                              \
          dpq += V * dp[f][j]; duq += V * du[f][j]; dvq += V * dv[f][j];                 \
          dwq += V * dw[f][j]; auq += V * au[f][j]; avq += V * av[f][j];                 \
          awq += V * aw[f][j];                                                           \
        }                                                                                \
        const R *g = fgeoq + (long)k * NFACES * 4 * Nfq + f * 4 * Nfq + n;               \
        const R nx = g[0], ny = g[Nfq], nz = g[2 * Nfq], sJ = g[3 * Nfq];                \
        const R Unjump = duq * nx + dvq * ny + dwq * nz;                                 \
        const R Unavg = auq * nx + avq * ny + awq * nz;                                  \
        const R pflux = -(Unavg - (R)0.5 * dpq) * sJ;                                    \
        const R Uflux = (R)-0.5 * (dpq - Unjump) * sJ;                                   \
        fp[f][n] = pflux; fu[f][n] = Uflux * nx; fv[f][n] = Uflux * ny; fw[f][n] = Uflux * nz; \
      }                                                                                  \
    R *o = rhsQ + (long)k * 4 * Np;                                                       \
    for (int i = 0; i < Np; ++i) {                                                       \
      R rp = o[i], ru = o[Np + i], rv = o[2 * Np + i], rw = o[3 * Np + i];               \
      for (int f = 0; f < NFACES; ++f)                                                   \
        for (int j = 0; j < Nfq; ++j) {                                                  \
          const R P = Pfq[i + (j + f * Nfq) * Np];                                       \
          rp += P * fp[f][j]; ru += P * fu[f][j]; rv += P * fv[f][j]; rw += P * fw[f][j]; \
        }                                                                                \
      o[i] = rp; o[Np + i] = ru; o[2 * Np + i] = rv; o[3 * Np + i] = rw;                 \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
void orc_volume_planar_##SFX(int Kl, const int *Klist, int Np, const R *vgeo, const R *J, \
                             const R *Dr, const R *Ds, const R *Dt, const R *Q, R *rhsQ)  \
{                                                                                        \
  _Pragma("omp parallel for schedule(static) if(Kl >= 1024)")                             \
  for (int kk = 0; kk < Kl; ++kk) {                                                      \
    const int k = Klist[kk];                                                             \
    const R *G = vgeo + (long)k * 9;                                                     \
    const R rx = G[0], ry = G[1], rz = G[2], sx = G[3], sy = G[4], sz = G[5];            \
    const R tx = G[6], ty = G[7], tz = G[8];                                             \
    R sUr[MAXN], sUs[MAXN], sUt[MAXN];                                                   \
    const R *q = Q + (long)k * 4 * Np;                                                    \
    for (int n = 0; n < Np; ++n) {                                                       \
      const R un = q[Np + n], vn = q[2 * Np + n], wn = q[3 * Np + n];                    \
      sUr[n] = un * rx + vn * ry + wn * rz;                                              \
      sUs[n] = un * sx + vn * sy + wn * sz;                                              \
      sUt[n] = un * tx + vn * ty + wn * tz;                                              \
    }                                                                                    \
    R *o = rhsQ + (long)k * 4 * Np;                                                       \
    const R Jk = J[k];                                                                   \
    for (int n = 0; n < Np; ++n) {                                                       \
      R dpdr = 0, dpds = 0, dpdt = 0, divU = 0;                                          \
      for (int m = 0; m < Np; ++m) {                                                     \
        const R a = Dr[n + m * Np], b = Ds[n + m * Np], c = Dt[n + m * Np];              \
        const R pm = q[m];                                                               \
        dpdr += a * pm; dpds += b * pm; dpdt += c * pm;                                  \
        divU += a * sUr[m]; divU += b * sUs[m]; divU += c * sUt[m];                      \
      }                                                                                  \
      const R dpdx = rx * dpdr + sx * dpds + tx * dpdt;                                  \
      const R dpdy = ry * dpdr + sy * dpds + ty * dpdt;                                  \
      const R dpdz = rz * dpdr + sz * dpds + tz * dpdt;                                  \
      o[n] = -divU * Jk; o[Np + n] = -dpdx * Jk; o[2 * Np + n] = -dpdy * Jk;             \
      o[3 * Np + n] = -dpdz * Jk;                                                        \
    }                                                                                    \
  }                                                                                      \
}                                                                                        \
                                                                                         \
void orc_surface_planar_##SFX(int Kl, const int *Klist, int Np, int Nfp, const R *fgeo,   \
                              const R *J, const int *Fmask, const int *vmapP,             \
                              const R *LIFT, const R *Q, R *rhsQ)                         \
{                                                                                        \
  _Pragma("omp parallel for schedule(static) if(Kl >= 1024)")                             \
  for (int kk = 0; kk < Kl; ++kk) {                                                      \
    const int k = Klist[kk];                                                             \
    R spf[NFACES * MAXN], suf[NFACES * MAXN], nxyz[3 * NFACES];                          \
    for (int n = 0; n < NFACES * Nfp; ++n) {                                             \
      const int f = n / Nfp;                                                             \
      int idM = Fmask[n] + k * Np * 4;                                                   \
      int idP = vmapP[n + (long)k * NFACES * Nfp];                                       \
      const int bnd = idM == idP;                                                        \
      const R *F = fgeo + (long)k * 16 + 4 * f;                                          \
      const R Fscale = F[0] * J[k], nx = F[1], ny = F[2], nz = F[3];                     \
      nxyz[3 * f] = nx; nxyz[3 * f + 1] = ny; nxyz[3 * f + 2] = nz;                      \
      const R pM = Q[idM], uM = Q[idM + Np], vM = Q[idM + 2 * Np], wM = Q[idM + 3 * Np]; \
      const R pP = Q[idP], uP = Q[idP + Np], vP = Q[idP + 2 * Np], wP = Q[idP + 3 * Np]; \
      R pjump = pP - pM;                                                                 \
      R Unjump = (uP - uM) * nx + (vP - vM) * ny + (wP - wM) * nz;                       \
      if (bnd) { pjump = -2 * pM; Unjump = 0; }                                          \
      spf[n] = (R)0.5 * (pjump - Unjump) * Fscale;                                       \
      suf[n] = (R)0.5 * (Unjump - pjump) * Fscale;                                       \
    }                                                                                    \
    R *o = rhsQ + (long)k * 4 * Np;                                                       \
    for (int n = 0; n < Np; ++n) {                                                       \
      R v1 = 0, v2 = 0, v3 = 0, v4 = 0;                                                  \
      for (int m = 0; m < NFACES * Nfp; ++m) {                                           \
        const R L = LIFT[n + m * Np];                                                    \
        const int fm = m / Nfp;                                                          \
        const R dfm = suf[m];                                                            \
        v1 += L * spf[m];                                                                \
        v2 += L * dfm * nxyz[3 * fm];                                                    \
        v3 += L * dfm * nxyz[1 + 3 * fm];                                                \
        v4 += L * dfm * nxyz[2 + 3 * fm];                                                \
      }                                                                                  \
      o[n] += v1; o[Np + n] += v2; o[2 * Np + n] += v3; o[3 * Np + n] += v4;             \
    }                                                                                    \
  }                                                                                      \
}

DEFINE_CURVED(double, f64)
DEFINE_CURVED(float, f32)
