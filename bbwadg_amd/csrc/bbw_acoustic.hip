// Acoustic BB-WADG kernels for gfx950 (p, u, v, w; Q[K][4][Np]).
//
// Plain path (one launch each, same data flow as the reference's stage):
//   k_volume       rhs  = (-div u, -grad p) from the 4-nnz barycentric operators
//                  (cpp/okl/WaveKernels.okl:344-474)
//   k_surface      rhs += EEL . blkdiag(L0) . upwind flux (cpp/okl/WaveKernels.okl:479-647)
//   k_update_wadg  Vq -> x c2/J -> Pq, LSRK stage (cpp/okl/WaveKernelsWADG.okl:1760-1887)
//   k_update       LSRK stage only (cpp/okl/WaveKernels.okl:262-313)
// Fused path:
//   k_fused        all of the above for a tile of elements in one launch per stage, reading
//                  Q_in and writing Q_out (ping-pong), so no stage-internal array (rhsQ)
//                  round-trips through HBM.
//
// All kernels stage whole element tiles through LDS with flat, fully coalesced loads (an
// element's Nf*Np coefficients are contiguous, and so is a tile of consecutive elements).
#include "bbw_internal.h"

namespace bbw {

__device__ __forceinline__ float src_time(const float *ftime_dev, float ftime) {
  return ftime_dev ? *ftime_dev : ftime;
}

// slot of element k in the ascending source element list, or -1
__device__ __forceinline__ int src_slot(const SrcArgs &s, int k) {
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

// ------------------------------------------------------------------ volume
template <int N>
__global__ __launch_bounds__(256) void k_volume(int K, int KB, const float *__restrict__ vgeo,
                                                const int4 *__restrict__ Dids,
                                                const float4 *__restrict__ Dvals,
                                                const float *__restrict__ Q,
                                                float *__restrict__ rhs) {
  constexpr int Np = Dims<N>::Np, Nf = 4;
  extern __shared__ float lds[];
  float *sQ = lds;               // [KB][4][Np]: p, then Ur, Us, Ut
  float *sG = lds + KB * Nf * Np;  // [KB][9]
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  const float *Qt = Q + (size_t)k0 * Nf * Np;
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) sQ[i] = Qt[i];
  for (int i = threadIdx.x; i < nk * 9; i += blockDim.x) sG[i] = vgeo[(size_t)k0 * 9 + i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    float *s = sQ + e * Nf * Np;
    const float *G = sG + 9 * e;
    const float u = s[Np + n], v = s[2 * Np + n], ww = s[3 * Np + n];
    s[Np + n] = u * G[0] + v * G[1] + ww * G[2];
    s[2 * Np + n] = u * G[3] + v * G[4] + ww * G[5];
    s[3 * Np + n] = u * G[6] + v * G[7] + ww * G[8];
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *sp = sQ + e * Nf * Np, *sUr = sp + Np, *sUs = sp + 2 * Np, *sUt = sp + 3 * Np;
    const float *G = sG + 9 * e;
    const float4 dv = Dvals[n];
    const int4 D1 = Dids[n], D2 = Dids[Np + n], D3 = Dids[2 * Np + n], D4 = Dids[3 * Np + n];
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
    const float divU = .5f * (dU1 - dU2);
    float *o = rhs + (size_t)(k0 + e) * Nf * Np + n;
    o[0] = -divU;
    o[Np] = -(G[0] * dpdr + G[3] * dpds + G[6] * dpdt);
    o[2 * Np] = -(G[1] * dpdr + G[4] * dpds + G[7] * dpdt);
    o[3 * Np] = -(G[2] * dpdr + G[5] * dpds + G[8] * dpdt);
  }
}

// ------------------------------------------------------------------ surface
template <int N>
__global__ __launch_bounds__(256) void k_surface(int K, int KB, const float *__restrict__ fgeo,
                                                 const int *__restrict__ Fmask,
                                                 const int *__restrict__ vmapP,
                                                 const int *__restrict__ EEL_ids,
                                                 const float *__restrict__ EEL_vals,
                                                 const int *__restrict__ L0_ids,
                                                 const float *__restrict__ L0_vals,
                                                 const float *__restrict__ Q,
                                                 float *__restrict__ rhs) {
  constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, Nf = 4;
  constexpr int EEL_nnz = Dims<N>::EEL_nnz, L0_nnz = Dims<N>::L0_nnz;
  extern __shared__ float lds[];
  float *sF = lds;                    // [KB][16]: per face Fscale nx ny nz
  float *sPf = sF + KB * 16;          // [KB][NfpNf]
  float *sUf = sPf + KB * NfpNf;
  float *sPt = sUf + KB * NfpNf;
  float *sUt = sPt + KB * NfpNf;
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  for (int i = threadIdx.x; i < nk * 16; i += blockDim.x) sF[i] = fgeo[(size_t)k0 * 16 + i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
    const int e = w / NfpNf, n = w - e * NfpNf, f = n / Nfp, k = k0 + e;
    const float *F = sF + 16 * e + 4 * f;
    const int idM = Fmask[n] + k * Np * Nf;
    const int idP = vmapP[(size_t)k * NfpNf + n];
    const float pM = Q[idM], uM = Q[idM + Np], vM = Q[idM + 2 * Np], wM = Q[idM + 3 * Np];
    const float pP = Q[idP], uP = Q[idP + Np], vP = Q[idP + 2 * Np], wP = Q[idP + 3 * Np];
    float pjump = pP - pM;
    float Unjump = (uP - uM) * F[1] + (vP - vM) * F[2] + (wP - wM) * F[3];
    if (idM == idP) {
      pjump = -2.f * pM;
      Unjump = 0.f;
    }
    sPf[w] = .5f * (pjump - Unjump) * F[0];
    sUf[w] = .5f * (Unjump - pjump) * F[0];
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
    const int e = w / NfpNf, n = w - e * NfpNf, f = n / Nfp, nt = n - f * Nfp;
    const float *pf = sPf + e * NfpNf + f * Nfp, *uf = sUf + e * NfpNf + f * Nfp;
    float v1 = 0.f, v2 = 0.f;
#pragma unroll
    for (int j = 0; j < L0_nnz; ++j) {
      const float L = L0_vals[nt + j * Nfp];
      const int id = L0_ids[nt + j * Nfp];
      v1 += L * pf[id];
      v2 += L * uf[id];
    }
    sPt[w] = v1;
    sUt[w] = v2;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *pt = sPt + e * NfpNf, *ut = sUt + e * NfpNf, *F = sF + 16 * e;
    float *o = rhs + (size_t)(k0 + e) * Nf * Np + n;
    float v1 = o[0], v2 = o[Np], v3 = o[2 * Np], v4 = o[3 * Np];
#pragma unroll
    for (int j = 0; j < EEL_nnz; ++j) {
      const int col = EEL_ids[n + j * Np];
      const float E = EEL_vals[n + j * Np];
      const int fc = col / Nfp;
      const float Uf = ut[col];
      v1 += E * pt[col];
      v2 += E * Uf * F[4 * fc + 1];
      v3 += E * Uf * F[4 * fc + 2];
      v4 += E * Uf * F[4 * fc + 3];
    }
    o[0] = v1;
    o[Np] = v2;
    o[2 * Np] = v3;
    o[3 * Np] = v4;
  }
}

// ------------------------------------------------------------------ WADG update
template <int N>
__global__ __launch_bounds__(256) void k_update_wadg(int K, int KB, const float *__restrict__ Vq,
                                                     const float *__restrict__ Pq,
                                                     const float *__restrict__ Jq,
                                                     const float *__restrict__ c2q, float fa,
                                                     float fb, float fdt, SrcArgs src,
                                                     const float *ftime_dev, float ftime,
                                                     const float *__restrict__ rhs,
                                                     float *__restrict__ res,
                                                     float *__restrict__ Q) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq, Nf = 4;
  extern __shared__ float lds[];
  float *sR = lds;                 // [KB][4][Np]
  float *sQq = lds + KB * Nf * Np;  // [KB][4][Nq]
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);
  const float *Rt = rhs + (size_t)k0 * Nf * Np;
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) sR[i] = Rt[i];
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Nq; w += blockDim.x) {
    const int e = w / Nq, q = w - e * Nq, k = k0 + e;
    const float *r = sR + e * Nf * Np;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll 8
    for (int j = 0; j < Np; ++j) {
      const float v = Vq[q + j * Nq];
      a0 += v * r[j];
      a1 += v * r[Np + j];
      a2 += v * r[2 * Np + j];
      a3 += v * r[3 * Np + j];
    }
    const float c2 = c2q[(size_t)k * Nq + q];
    const float invJ = Jq ? 1.f / Jq[(size_t)k * Nq + q] : 1.f;
    float *s = sQq + e * Nf * Nq;
    s[q] = a0 * invJ * c2;
    s[Nq + q] = a1 * invJ;
    s[2 * Nq + q] = a2 * invJ;
    s[3 * Nq + q] = a3 * invJ;
  }
  __syncthreads();
  const float ft = src_time(ftime_dev, ftime);
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np, k = k0 + e;
    const float *s = sQq + e * Nf * Nq;
    float r0 = 0.f, r1 = 0.f, r2 = 0.f, r3 = 0.f;
#pragma unroll 8
    for (int j = 0; j < Nq; ++j) {
      const float p = Pq[n + j * Np];
      r0 += p * s[j];
      r1 += p * s[Nq + j];
      r2 += p * s[2 * Nq + j];
      r3 += p * s[3 * Nq + j];
    }
    const int slot = src_slot(src, k);
    if (slot >= 0) {
      const float c = ft * src.coef[(size_t)slot * Np + n];
      r0 += src.w[0] * c;
      r1 += src.w[1] * c;
      r2 += src.w[2] * c;
      r3 += src.w[3] * c;
    }
    const size_t id = (size_t)k * Nf * Np + n;
    float x;
    x = fa * res[id] + fdt * r0;            res[id] = x;            Q[id] += fb * x;
    x = fa * res[id + Np] + fdt * r1;       res[id + Np] = x;       Q[id + Np] += fb * x;
    x = fa * res[id + 2 * Np] + fdt * r2;   res[id + 2 * Np] = x;   Q[id + 2 * Np] += fb * x;
    x = fa * res[id + 3 * Np] + fdt * r3;   res[id + 3 * Np] = x;   Q[id + 3 * Np] += fb * x;
  }
}

__global__ __launch_bounds__(256) void k_update(size_t n, float fa, float fb, float fdt,
                                                const float *__restrict__ rhs,
                                                float *__restrict__ res, float *__restrict__ Q) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float r = fa * res[i] + fdt * rhs[i];
    res[i] = r;
    Q[i] += fb * r;
  }
}

// ------------------------------------------------------------------ fused stage
// One tile of KB elements: Q_in tile -> LDS; volume and upwind flux; L0 then EEL lift; WADG
// projection of p (and of u, v, w when PROJ_U); source; LSRK; Q_out = Q_in + fb*res.
template <int N, bool PROJ_U>
__global__ __launch_bounds__(256) void k_fused(int K, int KB, const float *__restrict__ vgeo,
                                               const float *__restrict__ fgeo,
                                               const int *__restrict__ vmapP,
                                               const int *__restrict__ Fmask,
                                               const int4 *__restrict__ Dids,
                                               const float4 *__restrict__ Dvals,
                                               const int *__restrict__ EEL_ids,
                                               const float *__restrict__ EEL_vals,
                                               const int *__restrict__ L0_ids,
                                               const float *__restrict__ L0_vals,
                                               const float *__restrict__ Vq,
                                               const float *__restrict__ Pq,
                                               const float *__restrict__ c2q, float fa, float fb,
                                               float fdt, SrcArgs src, const float *ftime_dev,
                                               float ftime, const float *__restrict__ Qin,
                                               float *__restrict__ res,
                                               float *__restrict__ Qout) {
  constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, Nq = Dims<N>::Nq;
  constexpr int Nf = 4, EEL_nnz = Dims<N>::EEL_nnz, L0_nnz = Dims<N>::L0_nnz;
  constexpr int NP = PROJ_U ? 4 : 1;                 // projected fields
  constexpr int FLX = 4 * NfpNf > NP * Nq ? 4 * NfpNf : NP * Nq;
  extern __shared__ float lds[];
  float *sQ = lds;                     // [KB][4][Np]  Q_in
  float *sU = sQ + KB * Nf * Np;       // [KB][3][Np]  contravariant velocity
  float *sR = sU + KB * 3 * Np;        // [KB][4][Np]  rhs
  float *sG = sR + KB * Nf * Np;       // [KB][9]
  float *sF = sG + KB * 9;             // [KB][16]
  float *sX = sF + KB * 16;            // [KB][FLX]    flux work, then cubature values
  const int k0 = blockIdx.x * KB;
  const int nk = min(KB, K - k0);

  const float *Qt = Qin + (size_t)k0 * Nf * Np;
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) sQ[i] = Qt[i];
  for (int i = threadIdx.x; i < nk * 9; i += blockDim.x) sG[i] = vgeo[(size_t)k0 * 9 + i];
  for (int i = threadIdx.x; i < nk * 16; i += blockDim.x) sF[i] = fgeo[(size_t)k0 * 16 + i];
  __syncthreads();

  // contravariant velocity Ur, Us, Ut
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *s = sQ + e * Nf * Np;
    const float *G = sG + 9 * e;
    float *u = sU + e * 3 * Np;
    const float uu = s[Np + n], vv = s[2 * Np + n], ww = s[3 * Np + n];
    u[n] = uu * G[0] + vv * G[1] + ww * G[2];
    u[Np + n] = uu * G[3] + vv * G[4] + ww * G[5];
    u[2 * Np + n] = uu * G[6] + vv * G[7] + ww * G[8];
  }
  // upwind flux at face nodes (neighbour traces gathered from Q_in)
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
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
    float *X = sX + e * FLX;
    X[n] = .5f * (pjump - Unjump) * F[0];
    X[NfpNf + n] = .5f * (Unjump - pjump) * F[0];
  }
  __syncthreads();

  // volume terms -> sR ; L0 per face -> X[2NfpNf..4NfpNf)
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *sp = sQ + e * Nf * Np;
    const float *sUr = sU + e * 3 * Np, *sUs = sUr + Np, *sUt = sUr + 2 * Np;
    const float *G = sG + 9 * e;
    const float4 dv = Dvals[n];
    const int4 D1 = Dids[n], D2 = Dids[Np + n], D3 = Dids[2 * Np + n], D4 = Dids[3 * Np + n];
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
    float *r = sR + e * Nf * Np;
    r[n] = -.5f * (dU1 - dU2);
    r[Np + n] = -(G[0] * dpdr + G[3] * dpds + G[6] * dpdt);
    r[2 * Np + n] = -(G[1] * dpdr + G[4] * dpds + G[7] * dpdt);
    r[3 * Np + n] = -(G[2] * dpdr + G[5] * dpds + G[8] * dpdt);
  }
  for (int w = threadIdx.x; w < nk * NfpNf; w += blockDim.x) {
    const int e = w / NfpNf, n = w - e * NfpNf, f = n / Nfp, nt = n - f * Nfp;
    float *X = sX + e * FLX;
    const float *pf = X + f * Nfp, *uf = X + NfpNf + f * Nfp;
    float v1 = 0.f, v2 = 0.f;
#pragma unroll
    for (int j = 0; j < L0_nnz; ++j) {
      const float L = L0_vals[nt + j * Nfp];
      const int id = L0_ids[nt + j * Nfp];
      v1 += L * pf[id];
      v2 += L * uf[id];
    }
    X[2 * NfpNf + n] = v1;
    X[3 * NfpNf + n] = v2;
  }
  __syncthreads();

  // EEL lift accumulated into sR
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *pt = sX + e * FLX + 2 * NfpNf, *ut = pt + NfpNf, *F = sF + 16 * e;
    float *r = sR + e * Nf * Np;
    float v1 = r[n], v2 = r[Np + n], v3 = r[2 * Np + n], v4 = r[3 * Np + n];
#pragma unroll
    for (int j = 0; j < EEL_nnz; ++j) {
      const int col = EEL_ids[n + j * Np];
      const float E = EEL_vals[n + j * Np];
      const int fc = col / Nfp;
      const float Uf = ut[col];
      v1 += E * pt[col];
      v2 += E * Uf * F[4 * fc + 1];
      v3 += E * Uf * F[4 * fc + 2];
      v4 += E * Uf * F[4 * fc + 3];
    }
    r[n] = v1;
    r[Np + n] = v2;
    r[2 * Np + n] = v3;
    r[3 * Np + n] = v4;
  }
  __syncthreads();

  // WADG: interpolate to cubature, weight p by c2 (flux work area is free now)
  for (int w = threadIdx.x; w < nk * Nq; w += blockDim.x) {
    const int e = w / Nq, q = w - e * Nq, k = k0 + e;
    const float *r = sR + e * Nf * Np;
    float a[NP];
#pragma unroll
    for (int f = 0; f < NP; ++f) a[f] = 0.f;
#pragma unroll 8
    for (int j = 0; j < Np; ++j) {
      const float v = Vq[q + j * Nq];
#pragma unroll
      for (int f = 0; f < NP; ++f) a[f] += v * r[f * Np + j];
    }
    float *X = sX + e * FLX;
    X[q] = a[0] * c2q[(size_t)k * Nq + q];
#pragma unroll
    for (int f = 1; f < NP; ++f) X[f * Nq + q] = a[f];
  }
  __syncthreads();

  // project back, add source; result into sR
  const float ft = src_time(ftime_dev, ftime);
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np, k = k0 + e;
    const float *X = sX + e * FLX;
    float a[NP];
#pragma unroll
    for (int f = 0; f < NP; ++f) a[f] = 0.f;
#pragma unroll 8
    for (int j = 0; j < Nq; ++j) {
      const float p = Pq[n + j * Np];
#pragma unroll
      for (int f = 0; f < NP; ++f) a[f] += p * X[f * Nq + j];
    }
    float *r = sR + e * Nf * Np;
    float out[4];
    out[0] = a[0];
#pragma unroll
    for (int f = 1; f < 4; ++f) out[f] = PROJ_U ? a[f < NP ? f : 0] : r[f * Np + n];
    const int slot = src_slot(src, k);
    if (slot >= 0) {
      const float c = ft * src.coef[(size_t)slot * Np + n];
#pragma unroll
      for (int f = 0; f < 4; ++f) out[f] += src.w[f] * c;
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) r[f * Np + n] = out[f];
  }
  __syncthreads();

  // LSRK stage: res = fa*res + fdt*rhs ; Q_out = Q_in + fb*res (flat, coalesced)
  const size_t base = (size_t)k0 * Nf * Np;
  for (int i = threadIdx.x; i < nk * Nf * Np; i += blockDim.x) {
    const float x = fa * res[base + i] + fdt * sR[i];
    res[base + i] = x;
    Qout[base + i] = sQ[i] + fb * x;
  }
}

__global__ void k_record(int nrec, int Np, int NfNp, const int *__restrict__ elems,
                         const float *__restrict__ basis, const float *__restrict__ Q,
                         float *__restrict__ row) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const float *Qk = Q + (size_t)elems[r] * NfNp;
  float acc = 0.f;
  for (int j = 0; j < Np; ++j) acc += basis[(size_t)r * Np + j] * Qk[j];
  row[r] = acc;
}

// ------------------------------------------------------------------ launchers
static inline int pick_kb(int req, int dflt) { return req > 0 ? req : dflt; }

static SrcArgs make_src(const bbw_ctx *c) {
  SrcArgs s = ctx_src(c);
  return s;
}

template <int N>
static hipError_t volume_n(bbw_ctx *c) {
  constexpr int Np = Dims<N>::Np;
  const int KB = pick_kb(c->KblkV, (256 + Np - 1) / Np);
  const size_t sh = sizeof(float) * (size_t)KB * (4 * Np + 9);
  const int grid = (c->K + KB - 1) / KB;
  hipLaunchKernelGGL(k_volume<N>, dim3(grid), dim3(256), sh, c->stream, c->K, KB, c->vgeo,
                     c->Dids, c->Dvals, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

template <int N>
static hipError_t surface_n(bbw_ctx *c) {
  constexpr int NfpNf = Dims<N>::NfpNf;
  const int KB = pick_kb(c->KblkS, (256 + NfpNf - 1) / NfpNf);
  const size_t sh = sizeof(float) * (size_t)KB * (16 + 4 * NfpNf);
  const int grid = (c->K + KB - 1) / KB;
  hipLaunchKernelGGL(k_surface<N>, dim3(grid), dim3(256), sh, c->stream, c->K, KB, c->fgeo,
                     c->Fmask, c->vmapP, c->EEL_ids, c->EEL_vals, c->L0_ids, c->L0_vals,
                     c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

template <int N>
static hipError_t update_wadg_n(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                                float ft) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq;
  const int KB = pick_kb(c->KblkU, (256 + Nq - 1) / Nq);
  const size_t sh = sizeof(float) * (size_t)KB * 4 * (Np + Nq);
  const int grid = (c->K + KB - 1) / KB;
  hipLaunchKernelGGL(k_update_wadg<N>, dim3(grid), dim3(256), sh, c->stream, c->K, KB, c->Vq,
                     c->Pq, c->Jq, c->c2q, fa, fb, fdt, make_src(c), ftd, ft, c->rhs, c->res,
                     c->Q[c->cur]);
  return hipGetLastError();
}

template <int N>
static hipError_t fused_n(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                          float fdt, const float *ftd, float ft, int projU) {
  constexpr int Np = Dims<N>::Np, NfpNf = Dims<N>::NfpNf, Nq = Dims<N>::Nq;
  const int KB = pick_kb(c->KblkF, N >= 5 ? 8 : N >= 4 ? 8 : 16);
  const int NP = projU ? 4 : 1;
  const int FLX = 4 * NfpNf > NP * Nq ? 4 * NfpNf : NP * Nq;
  const size_t sh = sizeof(float) * (size_t)KB * (4 * Np + 3 * Np + 4 * Np + 9 + 16 + FLX);
  const int grid = (c->K + KB - 1) / KB;
  if (projU)
    hipLaunchKernelGGL((k_fused<N, true>), dim3(grid), dim3(256), sh, c->stream, c->K, KB,
                       c->vgeo, c->fgeo, c->vmapP, c->Fmask, c->Dids, c->Dvals, c->EEL_ids,
                       c->EEL_vals, c->L0_ids, c->L0_vals, c->Vq, c->Pq, c->c2q, fa, fb, fdt,
                       make_src(c), ftd, ft, Qin, c->res, Qout);
  else
    hipLaunchKernelGGL((k_fused<N, false>), dim3(grid), dim3(256), sh, c->stream, c->K, KB,
                       c->vgeo, c->fgeo, c->vmapP, c->Fmask, c->Dids, c->Dvals, c->EEL_ids,
                       c->EEL_vals, c->L0_ids, c->L0_vals, c->Vq, c->Pq, c->c2q, fa, fb, fdt,
                       make_src(c), ftd, ft, Qin, c->res, Qout);
  return hipGetLastError();
}

bool order_supported(int N) { return N >= 1 && N <= 9; }

hipError_t launch_volume(bbw_ctx *c) {
  if (c->plain_impl == BBW_PLAIN_TILED) return launch_volume_tiled(c);
  switch (c->N) {
    case 1: return volume_n<1>(c);
    case 2: return volume_n<2>(c);
    case 3: return volume_n<3>(c);
    case 4: return volume_n<4>(c);
    case 5: return volume_n<5>(c);
    case 6: return volume_n<6>(c);
    case 7: return volume_n<7>(c);
    case 8: return volume_n<8>(c);
    case 9: return volume_n<9>(c);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_surface(bbw_ctx *c) {
  if (c->plain_impl == BBW_PLAIN_TILED) return launch_surface_tiled(c);
  if (c->plain_impl == BBW_PLAIN_SLICE) return launch_surface_slice(c);
  switch (c->N) {
    case 1: return surface_n<1>(c);
    case 2: return surface_n<2>(c);
    case 3: return surface_n<3>(c);
    case 4: return surface_n<4>(c);
    case 5: return surface_n<5>(c);
    case 6: return surface_n<6>(c);
    case 7: return surface_n<7>(c);
    case 8: return surface_n<8>(c);
    case 9: return surface_n<9>(c);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_update_wadg(bbw_ctx *c, float fa, float fb, float fdt, const float *ftd,
                              float ft) {
  // orders 7..9 (no tiled kernels): the chunked MFMA update unless the flat kernels are asked for
  if (c->plain_impl == BBW_PLAIN_TILED || (c->N > BBW_MAX_ORDER_TILED && c->plain_impl == BBW_PLAIN_SLICE))
    return launch_update_wadg_mfma(c, fa, fb, fdt, ftd, ft);
  switch (c->N) {
    case 1: return update_wadg_n<1>(c, fa, fb, fdt, ftd, ft);
    case 2: return update_wadg_n<2>(c, fa, fb, fdt, ftd, ft);
    case 3: return update_wadg_n<3>(c, fa, fb, fdt, ftd, ft);
    case 4: return update_wadg_n<4>(c, fa, fb, fdt, ftd, ft);
    case 5: return update_wadg_n<5>(c, fa, fb, fdt, ftd, ft);
    case 6: return update_wadg_n<6>(c, fa, fb, fdt, ftd, ft);
    case 7: return update_wadg_n<7>(c, fa, fb, fdt, ftd, ft);
    case 8: return update_wadg_n<8>(c, fa, fb, fdt, ftd, ft);
    case 9: return update_wadg_n<9>(c, fa, fb, fdt, ftd, ft);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_update(bbw_ctx *c, float fa, float fb, float fdt) {
  const size_t n = (size_t)c->K * c->Nf * c->Np;
  const int grid = (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
  hipLaunchKernelGGL(k_update, dim3(grid), dim3(256), 0, c->stream, n, fa, fb, fdt, c->rhs,
                     c->res, c->Q[c->cur]);
  return hipGetLastError();
}

hipError_t launch_fused(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb, float fdt,
                        const float *ftd, float ft, int projU) {
  switch (c->N) {
    case 1: return fused_n<1>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU);
    case 2: return fused_n<2>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU);
    case 3: return fused_n<3>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU);
    case 4: return fused_n<4>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU);
    case 5: return fused_n<5>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU);
    case 6: return fused_n<6>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_record(bbw_ctx *c, const float *Q, float *row) {
  const int grid = (c->nrec + 127) / 128;
  hipLaunchKernelGGL(k_record, dim3(grid), dim3(128), 0, c->stream, c->nrec, c->Np,
                     c->Nf * c->Np, c->rec_elems, c->rec_basis, Q + (size_t)c->rec_field * c->Np, row);
  return hipGetLastError();
}

}  // namespace bbw
