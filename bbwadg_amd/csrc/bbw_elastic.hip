// Elastic velocity-stress BB-WADG kernels for gfx950 (9 fields: v1 v2 v3 sxx syy szz syz sxz sxy;
// Q[K][9][Np]).
//
//   k_volume_elas   rhs(v) = div sigma, rhs(sigma) = strain rates (elas_cpp/okl/ElasKernelsWADG.okl
//                   :25-124), derivatives from the 4-nnz barycentric BB operators instead of
//                   dense nodal Dr/Ds/Dt
//   k_surface_elas  central flux + tau_v/tau_s penalties, Q+ = 0 on the boundary (:128-284),
//                   lifted with EEL . blkdiag(L0) instead of the dense nodal LIFT
//   k_update_elas   Vq -> (rhoinv v, C : strain) -> Pq, + source, LSRK (:286-433)
#include "bbw_internal.h"

namespace bbw {

template <int N>
__global__ __launch_bounds__(256) void k_volume_elas(int K, int KB, const float *__restrict__ vgeo,
                                                     const int4 *__restrict__ Dids,
                                                     const float4 *__restrict__ Dvals,
                                                     const float *__restrict__ Q,
                                                     float *__restrict__ rhs) {
  constexpr int Np = Dims<N>::Np, Nf = 9;
  extern __shared__ float lds[];
  float *sQ = lds;                 // [KB][9][Np]
  float *sG = lds + KB * Nf * Np;  // [KB][9]
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  const float *Qt = Q + (size_t)k0 * Nf * Np;
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) sQ[i] = Qt[i];
  for (int i = threadIdx.x; i < nk * 9; i += blockDim.x) sG[i] = vgeo[(size_t)k0 * 9 + i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *G = sG + 9 * e;
    const float4 dv = Dvals[n];
    const int4 D1 = Dids[n], D2 = Dids[Np + n], D3 = Dids[2 * Np + n], D4 = Dids[3 * Np + n];
    float Qx[9], Qy[9], Qz[9];
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) {
      const float *s = sQ + (e * Nf + fl) * Np;
      const float p1 = dv.x * s[D1.x] + dv.y * s[D2.x] + dv.z * s[D3.x] + dv.w * s[D4.x];
      const float p2 = dv.x * s[D1.y] + dv.y * s[D2.y] + dv.z * s[D3.y] + dv.w * s[D4.y];
      const float p3 = dv.x * s[D1.z] + dv.y * s[D2.z] + dv.z * s[D3.z] + dv.w * s[D4.z];
      const float p4 = dv.x * s[D1.w] + dv.y * s[D2.w] + dv.z * s[D3.w] + dv.w * s[D4.w];
      const float dr = .5f * (p2 - p1), ds = .5f * (p3 - p1), dt = .5f * (p4 - p1);
      Qx[fl] = G[0] * dr + G[3] * ds + G[6] * dt;
      Qy[fl] = G[1] * dr + G[4] * ds + G[7] * dt;
      Qz[fl] = G[2] * dr + G[5] * ds + G[8] * dt;
    }
    float *o = rhs + (size_t)(k0 + e) * Nf * Np + n;
    o[0 * Np] = Qx[3] + Qy[8] + Qz[7];
    o[1 * Np] = Qx[8] + Qy[4] + Qz[6];
    o[2 * Np] = Qx[7] + Qy[6] + Qz[5];
    o[3 * Np] = Qx[0];
    o[4 * Np] = Qy[1];
    o[5 * Np] = Qz[2];
    o[6 * Np] = Qy[2] + Qz[1];
    o[7 * Np] = Qx[2] + Qz[0];
    o[8 * Np] = Qx[1] + Qy[0];
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_surface_elas(
    int K, int KB, const float *__restrict__ fgeo, const int *__restrict__ Fmask,
    const int *__restrict__ vmapP, const int *__restrict__ EEL_ids,
    const float *__restrict__ EEL_vals, const int *__restrict__ L0_ids,
    const float *__restrict__ L0_vals, float tau_v, float tau_s, const float *__restrict__ Q,
    float *__restrict__ rhs) {
  constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, Nf = 9;
  constexpr int EEL_nnz = Dims<N>::EEL_nnz, L0_nnz = Dims<N>::L0_nnz;
  extern __shared__ float lds[];
  float *sF = lds;                     // [KB][16]
  float *sFl = sF + KB * 16;           // [KB][9][NfpNf] flux
  float *sT = sFl + KB * Nf * NfpNf;   // [KB][9][NfpNf] L0 . flux
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  for (int i = threadIdx.x; i < nk * 16; i += blockDim.x) sF[i] = fgeo[(size_t)k0 * 16 + i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
    const int e = w / NfpNf, i = w - e * NfpNf, f = i / Nfp, k = k0 + e;
    const float *F = sF + 16 * e + 4 * f;
    const float Fscale = F[0], nx = F[1], ny = F[2], nz = F[3];
    int idM = Fmask[i] + k * Np * Nf;
    int idP = vmapP[(size_t)k * NfpNf + i];
    const bool isB = idM == idP;
    float dQ[9];
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) {
      dQ[fl] = -Q[idM];
      if (!isB) dQ[fl] += Q[idP];
      idM += Np;
      idP += Np;
    }
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
    float fl_[9];
    fl_[0] = fc[0] + tau_v * (fc[3] * nx + fc[8] * ny + fc[7] * nz);
    fl_[1] = fc[1] + tau_v * (fc[8] * nx + fc[4] * ny + fc[6] * nz);
    fl_[2] = fc[2] + tau_v * (fc[7] * nx + fc[6] * ny + fc[5] * nz);
    fl_[3] = fc[3] + tau_s * (fc[0] * nx);
    fl_[4] = fc[4] + tau_s * (fc[1] * ny);
    fl_[5] = fc[5] + tau_s * (fc[2] * nz);
    fl_[6] = fc[6] + tau_s * (fc[2] * ny + fc[1] * nz);
    fl_[7] = fc[7] + tau_s * (fc[2] * nx + fc[0] * nz);
    fl_[8] = fc[8] + tau_s * (fc[1] * nx + fc[0] * ny);
    float *X = sFl + e * Nf * NfpNf;
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) X[fl * NfpNf + i] = fl_[fl] * .5f * Fscale;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
    const int e = w / NfpNf, i = w - e * NfpNf, f = i / Nfp, nt = i - f * Nfp;
    const float *X = sFl + e * Nf * NfpNf + f * Nfp;
    float *T = sT + e * Nf * NfpNf;
    float v[9];
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) v[fl] = 0.f;
#pragma unroll
    for (int j = 0; j < L0_nnz; ++j) {
      const float L = L0_vals[nt + j * Nfp];
      const int id = L0_ids[nt + j * Nfp];
#pragma unroll
      for (int fl = 0; fl < Nf; ++fl) v[fl] += L * X[fl * NfpNf + id];
    }
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) T[fl * NfpNf + i] = v[fl];
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *T = sT + e * Nf * NfpNf;
    float v[9];
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) v[fl] = 0.f;
#pragma unroll
    for (int j = 0; j < EEL_nnz; ++j) {
      const int col = EEL_ids[n + j * Np];
      const float E = EEL_vals[n + j * Np];
#pragma unroll
      for (int fl = 0; fl < Nf; ++fl) v[fl] += E * T[fl * NfpNf + col];
    }
    float *o = rhs + (size_t)(k0 + e) * Nf * Np + n;
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) o[fl * Np] += v[fl];
  }
}

__device__ __forceinline__ int src_slot_e(const SrcArgs &s, int k) {
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

template <int N>
__global__ __launch_bounds__(256) void k_update_elas(
    int K, int KB, const float *__restrict__ Vq, const float *__restrict__ Pq,
    const float *__restrict__ rhoinvq, const float *__restrict__ lambdaq,
    const float *__restrict__ muq, SrcArgs src, const float *ftime_dev, float ftime, float fa,
    float fb, float fdt, const float *__restrict__ rhs, float *__restrict__ res,
    float *__restrict__ Q) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq, Nf = 9;
  extern __shared__ float lds[];
  float *sR = lds;                  // [KB][9][Np]
  float *sQq = lds + KB * Nf * Np;  // [KB][9][Nq]
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  const float *Rt = rhs + (size_t)k0 * Nf * Np;
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) sR[i] = Rt[i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Nq; w += blockDim.x) {
    const int e = w / Nq, q = w - e * Nq, k = k0 + e;
    const float *r = sR + e * Nf * Np;
    float a[9];
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) a[fl] = 0.f;
    for (int j = 0; j < Np; ++j) {
      const float v = Vq[q + j * Nq];
#pragma unroll
      for (int fl = 0; fl < Nf; ++fl) a[fl] += v * r[fl * Np + j];
    }
    const size_t id = (size_t)k * Nq + q;
    const float rho = rhoinvq[id], lam = lambdaq[id], mu = muq[id];
    const float A = 2.f * mu + lam;
    float *s = sQq + e * Nf * Nq;
    s[0 * Nq + q] = rho * a[0];
    s[1 * Nq + q] = rho * a[1];
    s[2 * Nq + q] = rho * a[2];
    s[3 * Nq + q] = A * a[3] + lam * a[4] + lam * a[5];
    s[4 * Nq + q] = A * a[4] + lam * a[3] + lam * a[5];
    s[5 * Nq + q] = A * a[5] + lam * a[3] + lam * a[4];
    s[6 * Nq + q] = mu * a[6];
    s[7 * Nq + q] = mu * a[7];
    s[8 * Nq + q] = mu * a[8];
  }
  __syncthreads();
  const float ft = ftime_dev ? *ftime_dev : ftime;
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np, k = k0 + e;
    const float *s = sQq + e * Nf * Nq;
    float a[9];
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) a[fl] = 0.f;
    for (int j = 0; j < Nq; ++j) {
      const float p = Pq[n + j * Np];
#pragma unroll
      for (int fl = 0; fl < Nf; ++fl) a[fl] += p * s[fl * Nq + j];
    }
    const int slot = src_slot_e(src, k);
    if (slot >= 0) {
      const float c = ft * src.coef[(size_t)slot * Np + n];
#pragma unroll
      for (int fl = 0; fl < Nf; ++fl) a[fl] += src.w[fl] * c;
    }
    const size_t id = (size_t)k * Nf * Np + n;
#pragma unroll
    for (int fl = 0; fl < Nf; ++fl) {
      const float x = fa * res[id + fl * Np] + fdt * a[fl];
      res[id + fl * Np] = x;
      Q[id + fl * Np] += fb * x;
    }
  }
}

static SrcArgs make_src_e(const bbw_ctx *c) {
  SrcArgs s = ctx_src(c);
  return s;
}

template <int N>
static hipError_t volume_elas_n(bbw_ctx *c) {
  constexpr int Np = Dims<N>::Np;
  const int KB = c->KblkV > 0 ? c->KblkV : (256 + Np - 1) / Np;
  const size_t sh = sizeof(float) * (size_t)KB * (9 * Np + 9);
  hipLaunchKernelGGL(k_volume_elas<N>, dim3((c->K + KB - 1) / KB), dim3(256), sh, c->stream, c->K,
                     KB, c->vgeo, c->Dids, c->Dvals, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

template <int N>
static hipError_t surface_elas_n(bbw_ctx *c) {
  constexpr int NfpNf = Dims<N>::NfpNf;
  const int KB = c->KblkS > 0 ? c->KblkS : (256 + NfpNf - 1) / NfpNf;
  const size_t sh = sizeof(float) * (size_t)KB * (16 + 18 * NfpNf);
  hipLaunchKernelGGL(k_surface_elas<N>, dim3((c->K + KB - 1) / KB), dim3(256), sh, c->stream,
                     c->K, KB, c->fgeo, c->Fmask, c->vmapP, c->EEL_ids, c->EEL_vals, c->L0_ids,
                     c->L0_vals, c->tau_v, c->tau_s, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

template <int N>
static hipError_t update_elas_n(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                float ft) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq;
  const int KB = c->KblkU > 0 ? c->KblkU : (256 + Nq - 1) / Nq;
  const size_t sh = sizeof(float) * (size_t)KB * 9 * (Np + Nq);
  hipLaunchKernelGGL(k_update_elas<N>, dim3((c->K + KB - 1) / KB), dim3(256), sh, c->stream, c->K,
                     KB, c->Vq, c->Pq, c->rhoinv, c->lambda, c->mu, make_src_e(c), ftd, ft, fa,
                     fb, fdt, c->rhs, c->res, c->Q[c->cur]);
  return hipGetLastError();
}

hipError_t launch_volume_elas(bbw_ctx *c) {
  if (c->plain_impl == BBW_PLAIN_TILED) return launch_volume_elas_tiled(c);
  switch (c->N) {
    case 1: return volume_elas_n<1>(c);
    case 2: return volume_elas_n<2>(c);
    case 3: return volume_elas_n<3>(c);
    case 4: return volume_elas_n<4>(c);
    case 5: return volume_elas_n<5>(c);
    case 6: return volume_elas_n<6>(c);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_surface_elas(bbw_ctx *c) {
  if (c->plain_impl == BBW_PLAIN_TILED) return launch_surface_elas_tiled(c);
  switch (c->N) {
    case 1: return surface_elas_n<1>(c);
    case 2: return surface_elas_n<2>(c);
    case 3: return surface_elas_n<3>(c);
    case 4: return surface_elas_n<4>(c);
    case 5: return surface_elas_n<5>(c);
    case 6: return surface_elas_n<6>(c);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_update_elas(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                              float ft) {
  if (c->plain_impl == BBW_PLAIN_TILED) return launch_update_elas_mfma(c, fa, fb, fdt, ftd, ft);
  switch (c->N) {
    case 1: return update_elas_n<1>(c, fa, fb, fdt, ftd, ft);
    case 2: return update_elas_n<2>(c, fa, fb, fdt, ftd, ft);
    case 3: return update_elas_n<3>(c, fa, fb, fdt, ftd, ft);
    case 4: return update_elas_n<4>(c, fa, fb, fdt, ftd, ft);
    case 5: return update_elas_n<5>(c, fa, fb, fdt, ftd, ft);
    case 6: return update_elas_n<6>(c, fa, fb, fdt, ftd, ft);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
