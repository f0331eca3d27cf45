// Surface term with the slice-by-slice Bernstein lift, for the acoustic plain stage
// (rk_surface_bern_slice, cpp/okl/WaveKernels.okl:652-939; selected for N > 6 by
// cpp/src/WaveOKL3d.cpp:1660-1665).
//
// The lift EEL . blkdiag(L0) of face f puts cEL[s] (E_{N-s}^N)^T (L0 flux_f) on slice s of
// that face (the nodes whose barycentric index for the opposite vertex is s).  (E_{N-s}^N)^T is
// s one-step degree reductions with 3 non-zeros per row, so the lift costs O(N^3) per face
// instead of the EEL form's O(N^4) (Nfp + 3 non-zeros per volume row).
//
// Per block of KB elements, in LDS:
//   1. Fscale, normals; upwind p- and U-flux per face node
//   2. slice 0 of each face = L0 flux (L0 applied per face)
//   3. slice s = reduction of slice s-1, s = 1..N, all faces and elements of the block at once
//      (one barrier per slice); every face keeps its own slice array, so no two threads ever
//      update one node (the reference accumulates all four faces into one array inside the
//      slice loop)
//   4. per volume node: rhs += sum_f cEL[s] lift_f at the node's row of face f's slices (row
//      and slice from slice_inv), the U-part times the face normal.
#include "bbw_internal.h"

namespace bbw {

template <int N>
__global__ __launch_bounds__(256) void k_surface_slice(int K, int KB, const float *__restrict__ fgeo,
                                                       const int *__restrict__ Fmask,
                                                       const int *__restrict__ vmapP,
                                                       const int *__restrict__ L0_ids,
                                                       const float *__restrict__ L0_vals,
                                                       const int4 *__restrict__ slice_inv,
                                                       const int *__restrict__ red_ids,
                                                       const float *__restrict__ red_vals,
                                                       const float *__restrict__ cEL,
                                                       const float *__restrict__ Q,
                                                       float *__restrict__ rhs) {
  constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf, Nf = 4;
  constexpr int L0_nnz = Dims<N>::L0_nnz;
  extern __shared__ float lds[];
  float *sF = lds;                    // [KB][16]: per face Fscale nx ny nz
  float *sPf = sF + KB * 16;          // [KB][NfpNf] p-flux
  float *sUf = sPf + KB * NfpNf;      // [KB][NfpNf] U-flux
  float *sLp = sUf + KB * NfpNf;      // [KB][4][Np] slices of each face (p)
  float *sLu = sLp + KB * 4 * Np;     // [KB][4][Np] slices of each face (U)
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
  // slice 0: L0 applied to each face's flux
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
    sLp[(e * 4 + f) * Np + nt] = v1;
    sLu[(e * 4 + f) * Np + nt] = v2;
  }
  // slices 1..N: slice s starts at row off = Np - Np(N - s) of a face's array, its reduction
  // rows at off - Nfp of red_*; it reads slice s - 1, which starts at Np - Np(N - s + 1)
  for (int s = 1; s <= N; ++s) {
    __syncthreads();
    const int m = N - s;
    const int nn = (m + 1) * (m + 2) / 2;
    const int off = Np - (m + 1) * (m + 2) * (m + 3) / 6;
    const int prev = Np - (m + 2) * (m + 3) * (m + 4) / 6;
    for (int w = threadIdx.x; w < nk * 4 * nn; w += blockDim.x) {
      const int ef = w / nn, r = w - ef * nn;
      const int row = off - Nfp + r;
      const int i0 = red_ids[3 * row], i1 = red_ids[3 * row + 1], i2 = red_ids[3 * row + 2];
      const float a0 = red_vals[3 * row], a1 = red_vals[3 * row + 1], a2 = red_vals[3 * row + 2];
      const float *lp = sLp + ef * Np + prev, *lu = sLu + ef * Np + prev;
      sLp[ef * Np + off + r] = a0 * lp[i0] + a1 * lp[i1] + a2 * lp[i2];
      sLu[ef * Np + off + r] = a0 * lu[i0] + a1 * lu[i1] + a2 * lu[i2];
    }
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nk * Np; w += blockDim.x) {
    const int e = w / Np, n = w - e * Np;
    const float *F = sF + 16 * e;
    const int4 iv = slice_inv[n];
    const int rows[4] = {iv.x, iv.y, iv.z, iv.w};
    float *o = rhs + (size_t)(k0 + e) * Nf * Np + n;
    float v1 = o[0], v2 = o[Np], v3 = o[2 * Np], v4 = o[3 * Np];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int sk = rows[f] & 0xffff, s = rows[f] >> 16;
      const float c = cEL[s];
      const float lp = c * sLp[(e * 4 + f) * Np + sk];
      const float lu = c * sLu[(e * 4 + f) * Np + sk];
      v1 += lp;
      v2 += lu * F[4 * f + 1];
      v3 += lu * F[4 * f + 2];
      v4 += lu * F[4 * f + 3];
    }
    o[0] = v1;
    o[Np] = v2;
    o[2 * Np] = v3;
    o[3 * Np] = v4;
  }
}

template <int N>
static hipError_t surface_slice_n(bbw_ctx *c) {
  constexpr int Np = Dims<N>::Np, NfpNf = Dims<N>::NfpNf;
  const int KB = c->KblkS > 0 ? c->KblkS : (256 + NfpNf - 1) / NfpNf;
  const size_t sh = sizeof(float) * (size_t)KB * (16 + 2 * NfpNf + 8 * Np);
  if (sh > 65536) return hipErrorInvalidConfiguration;
  const int grid = (c->K + KB - 1) / KB;
  hipLaunchKernelGGL(k_surface_slice<N>, dim3(grid), dim3(256), sh, c->stream, c->K, KB, c->fgeo,
                     c->Fmask, c->vmapP, c->L0_ids, c->L0_vals, c->slice_inv, c->red_ids,
                     c->red_vals, c->cEL, c->Q[c->cur], c->rhs);
  return hipGetLastError();
}

hipError_t launch_surface_slice(bbw_ctx *c) {
  switch (c->N) {
    case 1: return surface_slice_n<1>(c);
    case 2: return surface_slice_n<2>(c);
    case 3: return surface_slice_n<3>(c);
    case 4: return surface_slice_n<4>(c);
    case 5: return surface_slice_n<5>(c);
    case 6: return surface_slice_n<6>(c);
    case 7: return surface_slice_n<7>(c);
    case 8: return surface_slice_n<8>(c);
    case 9: return surface_slice_n<9>(c);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
