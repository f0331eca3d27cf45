// Discrete adjoint of the acoustic fused LSRK45 stage (gfx950).
//
// Forward stage (W = Pq diag(c2) Vq on p, identity on u, v, w; A = volume + surface):
//     r = W A Q + s(t);  res' = a res + dt r;  Q' = Q + b res'
// Reverse stage, given the adjoints lQ (of Q') and lres (of res'):
//     g = lres + b lQ;  lres <- a g;  z = W^T g;  grad_c2 += dt (Pq^T g_p) * y;
//     lQ <- lQ + dt A^T z,          y = Vq (A Q)_p at the stage input (saved by the forward)
// Kernels:
//   k_adj_wadg   g, lres update, t = Pq^T g_p, gradient, z_p = Vq^T (c2 t), z_u = g_u -> Z
//   k_adj_lift   h = 1/2 Fscale (L0^T EEL^T z_p - n . L0^T EEL^T z_u) per face node -> H
//   k_adj_apply  A^T z: volume transpose (D_m^T rows) + own and neighbour flux transposes
//                gathered through Fmask^-1 and mapP; lQ += dt A^T z
//   k_adj_inject receiver residuals into lQ_p (B_r^T (p_r - d_r))
#include "bbw_internal.h"

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
    float t = 0.f;
    for (int n = 0; n < Np; ++n) t += Pq[n + q * Np] * g[n];
    const size_t id = (size_t)(k0 + e) * Nq + q;
    grad[id] += fdt * t * Y[id];
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
                     c->Nf * c->Np, c->rec_elems, c->rec_basis, resid_row, c->lQ);
  return hipGetLastError();
}

}  // namespace bbw
